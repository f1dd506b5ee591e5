/*
 * fr_oracle.c -- CPU restatement of the HiFiLES FR residual + RK update.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never links this file.
 *
 * Every function cites the reference lines it restates.  Floating-point
 * operations are written in the same order as the reference CPU build with
 * -D_NO_BLAS (src/global.cpp:38-103 dgemm), so that, compiled with
 * -ffp-contract=off, the oracle reproduces the reference binary to the last bit
 * on the same mesh and state.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   gauss table, eval_lagrange/eval_d_lagrange, opp_0/opp_4  -- polynomial
 *       exactness (interpolation/differentiation of degree-P polynomials),
 *       Gauss nodes = roots of P_{n} to 1e-15, and reference residuals.
 *   eval_d_vcjh_1d / compute_eta / opp_3 -- DG limit (eta=0) equals the
 *       derivative of the right Radau polynomial; exactness of the corrected
 *       divergence for polynomial fluxes; reference residuals (SD scheme).
 *   set_transforms metrics -- closed forms on affine elements; reference residuals.
 *   fluxes (Euler, NS viscous, Rusanov, LDG) -- free-stream preservation,
 *       conservation, reference residuals from oracle/_ref on seeded states.
 *   RK45 -- exactness for polynomial-in-time ODEs (order 4), reference
 *       RK45 trajectories and the TGV 25-step residual norms of the reference's
 *       own regression test (testcases/regression_tests.py:300-307).
 */
#include "fr_oracle.h"
#include "tet_points.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* 1-D Gauss-Legendre points as tabulated (15 decimals) by the reference,
 * data/loc_1d_gauss_pts.dat:1-60 (orders 0..6).  The tabulated values are used
 * verbatim (not recomputed) because operator parity at 1e-15 depends on them. */
static const double gauss_tab[7][7] = {
  {0.0},
  {-0.577350269189626, 0.577350269189626},
  {-0.774596669241483, -0.000000000000000, 0.774596669241483},
  {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
  {-0.906179845938664, -0.538469310105683, -0.000000000000000, 0.538469310105683,
   0.906179845938664},
  {-0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197,
   0.661209386466265, 0.932469514203152},
  {-0.949107912342758, -0.741531185599395, -0.405845151377397, 0.000000000000000,
   0.405845151377397, 0.741531185599394, 0.949107912342758}};


/* ------------------------------------------------------------------------- */
/* 1-D basis functions, src/funcs.cpp:57-250 */

/* src/funcs.cpp:57-73 */
static double eval_lagrange(double r, int mode, const double* pts, int n) {
  double t = 1.0;
  for (int i = 0; i < n; i++)
    if (i != mode) t = t * ((r - pts[i]) / (pts[mode] - pts[i]));
  return t;
}

/* src/funcs.cpp:78-105 */
static double eval_d_lagrange(double r, int mode, const double* pts, int n) {
  double t0 = 0.0;
  for (int i = 0; i < n; i++) {
    if (i == mode) continue;
    double t1 = 1.0, t2 = 1.0;
    for (int j = 0; j < n; j++) {
      if (j != mode && j != i) t1 = t1 * (r - pts[j]);
      if (j != mode) t2 = t2 * (pts[mode] - pts[j]);
    }
    t0 = t0 + (t1 / t2);
  }
  return t0;
}

/* src/funcs.cpp:161-179 */
static double eval_legendre(double r, int mode) {
  if (mode == 0) return 1.0;
  if (mode == 1) return r;
  return ((2 * mode - 1) * r * eval_legendre(r, mode - 1) -
          (mode - 1) * eval_legendre(r, mode - 2)) / mode;
}

/* src/funcs.cpp:183-212 */
static double eval_d_legendre(double r, int mode) {
  double d = 0.0;
  if (mode == 0) return 0;
  if (r > -1.0 && r < 1.0) {
    d = (mode * ((r * eval_legendre(r, mode)) - eval_legendre(r, mode - 1))) / ((r * r) - 1.0);
  } else {
    if (r == -1.0) d = pow(-1.0, mode - 1.0) * 0.5 * mode * (mode + 1.0);
    if (r == 1.0) d = 0.5 * mode * (mode + 1.0);
  }
  return d;
}

/* src/funcs.cpp:216-250; mode 0 = left correction function, 1 = right */
static double eval_d_vcjh_1d(double r, int mode, int order, double eta) {
  double d = 0.0;
  if (mode == 0) {
    if (order == 0)
      d = 0.5 * pow(-1.0, order) *
          (eval_d_legendre(r, order) - ((eval_d_legendre(r, order + 1)) / (1.0 + eta)));
    else
      d = 0.5 * pow(-1.0, order) *
          (eval_d_legendre(r, order) -
           (((eta * eval_d_legendre(r, order - 1)) + eval_d_legendre(r, order + 1)) / (1.0 + eta)));
  } else {
    if (order == 0)
      d = 0.5 * (eval_d_legendre(r, order) + ((eval_d_legendre(r, order + 1)) / (1.0 + eta)));
    else
      d = 0.5 * (eval_d_legendre(r, order) +
                 (((eta * eval_d_legendre(r, order - 1)) + eval_d_legendre(r, order + 1)) /
                  (1.0 + eta)));
  }
  return d;
}

/* src/funcs.cpp:1503-1546 (schemes 1..3; scheme 4 (c+) is not restated) */
static double compute_eta(int scheme, int order) {
  if (scheme == 1) return 0.0;
  if (scheme == 2) return (1.0 * (order)) / (1.0 * (order + 1));
  if (scheme == 3) return (1.0 * (order + 1)) / (1.0 * order);
  fprintf(stderr, "fr_oracle: vcjh scheme %d not restated\n", scheme);
  abort();
}

/* src/global.cpp:38-103, the no-BLAS dgemm: C = alpha*A*B + beta*C, column-major,
 * accumulation order l = 0..Acols-1 per column.  Columns are independent, so the
 * OpenMP split over j is bit-identical to the serial loop. */
static void dgemm_ref(int Arows, int Bcols, int Acols, double alpha, double beta,
                      const double* a, const double* b, double* c) {
  if (Arows == 0 || Bcols == 0 || ((alpha == 0. || Acols == 0) && beta == 1.)) return;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < Bcols; j++) {
    double* cj = c + (long)j * Arows;
    const double* bj = b + (long)j * Acols;
    if (beta == 0.) {
      for (int i = 0; i < Arows; i++) cj[i] = 0.;
    } else if (beta != 1.) {
      for (int i = 0; i < Arows; i++) cj[i] = beta * cj[i];
    }
    for (int l = 0; l < Acols; l++) {
      double temp = alpha * bj[l];
      const double* al = a + (long)l * Arows;
      for (int i = 0; i < Arows; i++) cj[i] += temp * al[i];
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Solver state */

struct or_solver {
  or_deck deck;
  or_phys ph;
  int ele_type, order, nd, nf, nu, nfp, nfaces, nfpf, nsp, n_eles, n_verts;
  double loc1d[8];
  double *loc_upts;   /* (d,Nu) */
  double *tloc_fpts;  /* (d,Nfp) */
  double *tnorm_fpts; /* (d,Nfp) */
  double *opp_0, *opp_1[3], *opp_2[3], *opp_3, *opp_4[3], *opp_5[3], *opp_6;
  double *shape;      /* (d,nsp,N) */
  double *detjac_upts, *JGinv_upts, *tdA_fpts, *norm_fpts, *pos_upts, *pos_fpts;
  double *disu_upts[2], *div_tconf_upts, *disu_fpts, *tdisf_upts, *norm_tdisf_fpts,
      *norm_tconf_fpts, *delta_disu_fpts, *grad_disu_upts, *grad_disu_fpts, *src_upts;
  double *sgsf_upts, *sgsf_fpts; /* LES: transformed SGS flux (Nu,N,nf,d), (Nfp,N,nf,d) */
  double *sgs_nu;                /* SGS_model 10: kinematic eddy viscosity per point (Nu,N) */
  int n_inters;
  double* inv_vdm;           /* tets: inverse Vandermonde of the Dubiner basis (Nu,Nu) */
  int n_bdy;                 /* boundary flux points */
  int *bdy_fpt, *bdy_type;   /* (n_bdy) flux-point index fpt + Nfp*ele, bctype */
  int *face_l, *face_r; /* (n_inters*nfpf) flux-point indices fpt + Nfp*ele */
};

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "fr_oracle: out of memory\n"); abort(); }
  return p;
}

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* src/input.cpp:89-331 defaults */
void or_deck_defaults(or_deck* d) {
  memset(d, 0, sizeof(*d));
  d->equation = 0; d->viscous = 0; d->riemann_solve_type = 0; d->vis_riemann_solve_type = 0;
  d->ic_form = 1; d->order = 3; d->adv_type = 3; d->dt = 1e-3;
  d->tau = 0.; d->pen_fact = 0.5;
  d->gamma = 1.4; d->prandtl = .72; d->S_gas = 120.; d->T_gas = 291.15; d->R_gas = 286.9;
  d->mu_gas = 1.827E-5; d->fix_vis = 1.;
  d->Mach_free_stream = 0.1; d->nx_free_stream = 1.; d->ny_free_stream = 0.; d->nz_free_stream = 0.;
  d->Re_free_stream = 1600.; d->L_free_stream = 1.; d->T_free_stream = 300.; d->T_wall = 300.;
  d->Mach_wall = 0.; d->nx_wall = 0.; d->ny_wall = 0.; d->nz_wall = 0.;
  d->upts_type = 0; d->vcjh_scheme = 0; d->eta = 0.;
  d->LES = 0; d->SGS_model = 0; d->filter_ratio = 1.; d->const_src = 0.;
  d->dx_cyclic = INFINITY; d->dy_cyclic = INFINITY; d->dz_cyclic = INFINITY;
}

/* src/input.cpp:345-515 (the quantities the residual uses) */
void or_setup_params(const or_deck* d, or_phys* o) {
  memset(o, 0, sizeof(*o));
  o->gamma = d->gamma; o->prandtl = d->prandtl; o->fix_vis = d->fix_vis;
  o->tau = d->tau; o->pen_fact = d->pen_fact; o->const_src = d->const_src; o->dt = d->dt;
  o->T_wall = d->T_wall;
  if (d->viscous) {
    if (d->ic_form == 0) {   /* src/input.cpp:392-399 */
      o->fix_vis = 1.; o->R_ref = 1.; o->c_sth = 1.; o->rt_inf = 1.; o->mu_inf = 0.1;
    } else {
      double T_ref = d->T_free_stream, L_ref = d->L_free_stream;
      double uvw_ref = d->Mach_free_stream * sqrt(d->gamma * d->R_gas * d->T_free_stream);
      double mu_free_stream;
      if (d->fix_vis) mu_free_stream = d->mu_gas;
      else mu_free_stream = d->mu_gas * pow(d->T_free_stream / d->T_gas, 1.5) *
                            ((d->T_gas + d->S_gas) / (d->T_free_stream + d->S_gas));
      double rho_free_stream = (mu_free_stream * d->Re_free_stream) / (uvw_ref * d->L_free_stream);
      double rho_ref = rho_free_stream;
      double mu_ref = rho_ref * uvw_ref * L_ref;
      o->R_ref = (d->R_gas * T_ref) / (uvw_ref * uvw_ref);
      o->c_sth = d->S_gas / d->T_gas;
      o->mu_inf = d->mu_gas / mu_ref;
      o->rt_inf = d->T_gas * d->R_gas / (uvw_ref * uvw_ref);
      double uvw_wall = d->Mach_wall * sqrt(d->gamma * d->R_gas * d->T_wall);
      o->v_wall[0] = (uvw_wall * d->nx_wall) / uvw_ref;
      o->v_wall[1] = (uvw_wall * d->ny_wall) / uvw_ref;
      o->v_wall[2] = (uvw_wall * d->nz_wall) / uvw_ref;
      o->T_wall = d->T_wall / T_ref;
      o->prandtl_t = 0.9;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Tetrahedra: orthonormal (Dubiner) polynomial bases on the reference simplex
 * (r,s,t >= -1, r+s+t <= -1), src/funcs.cpp:1015-1480 */

/* (n-1)! for n >= 1, src/funcs.cpp:1096-1114 */
static double gamma_int(int n) {
  double g = 1.0;
  for (int i = 2; i < n; i++) g *= i;
  return g;
}

/* Jacobi polynomial P_n^(a,b)(x), normalised to unit norm on [-1,1] with weight
 * (1-x)^a (1+x)^b: the reference's three-term recursion with its operation order,
 * src/funcs.cpp:1117-1170 */
static double jacobi_n(double x, int a, int b, int n) {
  if (n == 0) {
    double t0 = pow(2.0, (-a - b - 1)), t1 = gamma_int(a + b + 2), t2 = gamma_int(a + 1) * gamma_int(b + 1);
    return sqrt(t0 * (t1 / t2));
  }
  if (n == 1) {
    double t0 = pow(2.0, (-a - b - 1)), t1 = gamma_int(a + b + 2), t2 = gamma_int(a + 1) * gamma_int(b + 1);
    double t3 = a + b + 3, t4 = (a + 1) * (b + 1), t5 = (x * (a + b + 2) + (a - b));
    return 0.5 * sqrt(t0 * (t1 / t2)) * sqrt(t3 / t4) * t5;
  }
  double t0 = n * (n + a + b) * (n + a) * (n + b);
  double t1 = ((2 * n) + a + b - 1) * ((2 * n) + a + b + 1);
  double t3 = (2 * n) + a + b;
  double t4 = (n - 1) * ((n - 1) + a + b) * ((n - 1) + a) * ((n - 1) + b);
  double t5 = ((2 * (n - 1)) + a + b - 1) * ((2 * (n - 1)) + a + b + 1);
  double t6 = (2 * (n - 1)) + a + b;
  double t7 = -((a * a) - (b * b));
  double t8 = ((2 * (n - 1)) + a + b) * ((2 * (n - 1)) + a + b + 2);
  double t9 = (2.0 / t3) * sqrt(t0 / t1);
  double t10 = (2.0 / t6) * sqrt(t4 / t5);
  double t11 = t7 / t8;
  double t12 = x * jacobi_n(x, a, b, n - 1);
  double t13 = t10 * jacobi_n(x, a, b, n - 2);
  double t14 = t11 * jacobi_n(x, a, b, n - 1);
  return (1.0 / t9) * (t12 - t13 - t14);
}

/* src/funcs.cpp:1174-1188 */
static double grad_jacobi_n(double x, int a, int b, int n) {
  return n == 0 ? 0.0 : sqrt(1.0 * n * (n + a + b + 1)) * jacobi_n(x, a + 1, b + 1, n - 1);
}

/* collapsed coordinates, src/funcs.cpp:1015-1031, 1067-1093 */
static void rs_to_ab(double r, double s, double* ab) {
  ab[0] = s == 1.0 ? -1.0 : 2.0 * ((1.0 + r) / (1.0 - s)) - 1.0;
  ab[1] = s;
}
static void rst_to_abc(double r, double s, double t, double* abc) {
  abc[0] = (s + t == 0.0) ? -1.0 : -2.0 * (1.0 + r) / (s + t) - 1.0;
  abc[1] = (t == 1.0) ? -1.0 : 2.0 * (1.0 + s) / (1.0 - t) - 1.0;
  abc[2] = t;
}

/* mode -> (i,j) of the triangle basis in the reference's order (k = i+j outer,
 * j inner); (i,j,k) of the tetrahedron basis (m = i+j+k, n = j+k, k inner) */
static void tri_mode(int mode, int* pi, int* pj) {
  int q = 0;
  for (int k = 0;; k++)
    for (int j = 0; j <= k; j++, q++)
      if (q == mode) { *pi = k - j; *pj = j; return; }
}
static void tet_mode(int mode, int* pi, int* pj, int* pk) {
  int q = 0;
  for (int m = 0;; m++)
    for (int n = 0; n <= m; n++)
      for (int k = 0; k <= n; k++, q++)
        if (q == mode) { *pk = k; *pj = n - k; *pi = m - n; return; }
}

/* orthonormal triangle basis, src/funcs.cpp:1190-1226 */
static double dubiner_2d(double r, double s, int mode) {
  int i, j;
  double ab[2];
  tri_mode(mode, &i, &j);
  rs_to_ab(r, s, ab);
  return sqrt(2.0) * jacobi_n(ab[0], 0, 0, i) * jacobi_n(ab[1], 2 * i + 1, 0, j) * pow(1.0 - ab[1], i);
}

/* orthonormal tetrahedron basis, src/funcs.cpp:1333-1378 */
static double dubiner_3d(double r, double s, double t, int mode) {
  int i, j, k;
  double c[3];
  tet_mode(mode, &i, &j, &k);
  rst_to_abc(r, s, t, c);
  return 2.0 * sqrt(2.0) * jacobi_n(c[0], 0, 0, i) * jacobi_n(c[1], 2 * i + 1, 0, j) *
         jacobi_n(c[2], 2 * i + 2 * j + 2, 0, k) * pow(1.0 - c[1], i) * pow(1.0 - c[2], i + j);
}

/* gradient of the tetrahedron basis by the chain rule through the collapsed
 * coordinates, src/funcs.cpp:1381-1480 */
static double grad_dubiner_3d(double r, double s, double t, int mode, int comp) {
  int i, j, k;
  double c[3];
  tet_mode(mode, &i, &j, &k);
  rst_to_abc(r, s, t, c);
  double fa = jacobi_n(c[0], 0, 0, i), gb = jacobi_n(c[1], 2 * i + 1, 0, j), hc = jacobi_n(c[2], 2 * (i + j) + 2, 0, k);
  double dfa = grad_jacobi_n(c[0], 0, 0, i), dgb = grad_jacobi_n(c[1], 2 * i + 1, 0, j),
         dhc = grad_jacobi_n(c[2], 2 * (i + j) + 2, 0, k);
  double hb = 0.5 * (1.0 - c[1]), hcc = 0.5 * (1.0 - c[2]);
  double dr = dfa * gb * hc;
  if (i > 0) dr *= pow(hb, i - 1);
  if (i + j > 0) dr *= pow(hcc, i + j - 1);
  double scale = pow(2.0, 2 * i + j + 1.5);
  if (comp == 0) return dr * scale;
  double ds = 0.5 * (1.0 + c[0]) * dr;
  double tmp = dgb * pow(hb, i);
  if (i > 0) tmp += (-0.5 * i) * gb * pow(hb, i - 1);
  if (i + j > 0) tmp *= pow(hcc, i + j - 1);
  tmp = fa * tmp * hc;
  ds += tmp;
  if (comp == 1) return ds * scale;
  double dt = 0.5 * (1.0 + c[0]) * dr + 0.5 * (1.0 + c[1]) * tmp;
  tmp = dhc * pow(hcc, i + j);
  if (i + j > 0) tmp -= 0.5 * (i + j) * (hc * pow(hcc, i + j - 1));
  tmp = fa * (gb * tmp) * pow(hb, i);
  dt += tmp;
  return dt * scale;
}

/* In-place inverse of an n x n column-major matrix: Gaussian elimination with full
 * pivoting applied to the identity, back substitution and the undoing of the
 * column swaps, in the order of inv_array (src/funcs.cpp:2719-2845). */
static void invert(double* A, int n) {
#define IN(i, j) A[(i) + n * (j)]
#define ID(i, j) I[(i) + n * (j)]
#define INV(i, j) X[(i) + n * (j)]
  double* I = xcalloc((size_t)n * n, sizeof(double));
  double* X = xcalloc((size_t)n * n, sizeof(double));
  int* sw0 = xcalloc((size_t)n, sizeof(int));
  int* sw1 = xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; i++) { sw0[i] = i; sw1[i] = i; ID(i, i) = 1.0; }
  for (int k = 0; k < n - 1; k++) {
    double mx = 0;
    int pi = k, pj = k;
    for (int i = k; i < n; i++)
      for (int j = k; j < n; j++) {
        double mag = IN(i, j) * IN(i, j);
        if (mag > mx) { pi = i; pj = j; mx = mag; }
      }
    int t = sw0[k]; sw0[k] = sw0[pi]; sw0[pi] = t;
    t = sw1[k]; sw1[k] = sw1[pj]; sw1[pj] = t;
    for (int i = 0; i < n; i++) { double a = IN(i, pj); IN(i, pj) = IN(i, k); IN(i, k) = a; }
    for (int j = 0; j < n; j++) {
      double a = IN(pi, j); IN(pi, j) = IN(k, j); IN(k, j) = a;
      a = ID(pi, j); ID(pi, j) = ID(k, j); ID(k, j) = a;
    }
    for (int i = k + 1; i < n; i++) {
      double first = IN(i, k);
      for (int j = 0; j < n; j++) {
        if (j >= k) IN(i, j) = IN(i, j) - ((first / IN(k, k)) * IN(k, j));
        ID(i, j) = ID(i, j) - ((first / IN(k, k)) * ID(k, j));
      }
    }
    for (int j = 0; j < k + 1; j++)
      for (int i = j + 1; i < n; i++) IN(i, j) = 0.0;
  }
  for (int i = n - 1; i >= 0; i = i - 1)
    for (int j = 0; j < n; j++) {
      double d = 0.0;
      for (int k = i + 1; k < n; k++) d = d + (IN(i, k) * INV(k, j));
      INV(i, j) = (ID(i, j) - d) / IN(i, i);
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[sw1[i] + n * j] = INV(i, j);
#undef IN
#undef ID
#undef INV
  free(I); free(X); free(sw0); free(sw1);
}

/* DG correction field of a tetrahedron flux point (eval_div_dg_tet,
 * src/eles_tets.cpp:1018-1131): the orthonormal projection onto P_P(tet) of the
 * face Lagrange polynomial of that flux point, integrated over its face with the
 * reference's 36-point triangle rule (face Jacobian sqrt(3) on the slanted face). */
static void tet_dg_lift(or_solver* s, double* opp3 /* (nu,nfp) */) {
  int nu = s->nu, nfpf = s->nfpf, nfp = s->nfp;
  double* Mf = xcalloc((size_t)nfpf * nfpf, sizeof(double));
  double* coef = xcalloc((size_t)nu, sizeof(double));
  for (int fpt = 0; fpt < nfp; fpt++) {
    int face = fpt / nfpf, ffp = fpt - face * nfpf;
    /* face coordinates of the face's flux points and the face Lagrange coefficients */
    for (int i = 0; i < nfpf; i++) {
      const double* t = s->tloc_fpts + 3 * (face * nfpf + i);
      double rf, sf;
      if (face == 0) { rf = t[0]; sf = t[2]; }
      else if (face == 1) { rf = t[2]; sf = t[1]; }
      else if (face == 2) { rf = t[0]; sf = t[2]; }
      else { rf = t[1]; sf = t[0]; }
      for (int j = 0; j < nfpf; j++) Mf[i + nfpf * j] = dubiner_2d(rf, sf, j);
    }
    invert(Mf, nfpf);  /* column ffp of the inverse: coefficients of the Lagrange polynomial */
    double face_jac = face == 0 ? sqrt(3.) : 1.0;
    for (int i = 0; i < nu; i++) {
      double integral = 0.;
      for (int q = 0; q < 36; q++) {
        double rf = tri_cub12[q][0], sf = tri_cub12[q][1];
        double r, ss, t;
        if (face == 0) { r = rf; t = sf; ss = -1. - t - r; }
        else if (face == 1) { r = -1.0; ss = sf; t = rf; }
        else if (face == 2) { r = rf; ss = -1.0; t = sf; }
        else { r = sf; ss = rf; t = -1.0; }
        double g = 0.;
        for (int k = 0; k < nfpf; k++) g += Mf[k + nfpf * ffp] * dubiner_2d(rf, sf, k);
        integral += tri_cub12[q][2] * dubiner_3d(r, ss, t, i) * g;
      }
      coef[i] = integral * face_jac;
    }
    for (int j = 0; j < nu; j++) {
      const double* x = s->loc_upts + 3 * j;
      double v = 0.;
      for (int i = 0; i < nu; i++) v += coef[i] * dubiner_3d(x[0], x[1], x[2], i);
      opp3[j + nu * fpt] = v;
    }
  }
  free(Mf); free(coef);
}

/* ------------------------------------------------------------------------- */
/* Element-type specific point sets and bases */

static int idx3(int n, int k, int j, int i) { return k + n * j + n * n * i; }

/* src/eles_hexas.cpp:917-928 / src/eles_quads.cpp:855-866 / src/eles_tets.cpp:805-822 */
static double nodal_basis(const or_solver* s, int idx, const double* loc) {
  int n = s->order + 1;
  if (s->ele_type == 2) {
    double v = 0.0;
    for (int i = 0; i < s->nu; i++) v += s->inv_vdm[i + s->nu * idx] * dubiner_3d(loc[0], loc[1], loc[2], i);
    return v;
  }
  if (s->ele_type == 4) {
    int i = idx / (n * n), j = (idx - n * n * i) / n, k = idx - n * j - n * n * i;
    return eval_lagrange(loc[0], k, s->loc1d, n) * eval_lagrange(loc[1], j, s->loc1d, n) *
           eval_lagrange(loc[2], i, s->loc1d, n);
  } else {
    int i = idx / n, j = idx - n * i;
    return eval_lagrange(loc[0], j, s->loc1d, n) * eval_lagrange(loc[1], i, s->loc1d, n);
  }
}

/* src/eles_hexas.cpp:945-973 / src/eles_quads.cpp:888-910 */
static double d_nodal_basis(const or_solver* s, int idx, int c, const double* loc) {
  int n = s->order + 1;
  const double* p = s->loc1d;
  if (s->ele_type == 2) {  /* src/eles_tets.cpp:843-860 */
    double v = 0.0;
    for (int i = 0; i < s->nu; i++) v += s->inv_vdm[i + s->nu * idx] * grad_dubiner_3d(loc[0], loc[1], loc[2], i, c);
    return v;
  }
  if (s->ele_type == 4) {
    int i = idx / (n * n), j = (idx - n * n * i) / n, k = idx - n * j - n * n * i;
    if (c == 0) return eval_d_lagrange(loc[0], k, p, n) * eval_lagrange(loc[1], j, p, n) * eval_lagrange(loc[2], i, p, n);
    if (c == 1) return eval_lagrange(loc[0], k, p, n) * eval_d_lagrange(loc[1], j, p, n) * eval_lagrange(loc[2], i, p, n);
    return eval_lagrange(loc[0], k, p, n) * eval_lagrange(loc[1], j, p, n) * eval_d_lagrange(loc[2], i, p, n);
  } else {
    int i = idx / n, j = idx - n * i;
    if (c == 0) return eval_d_lagrange(loc[0], j, p, n) * eval_lagrange(loc[1], i, p, n);
    return eval_lagrange(loc[0], j, p, n) * eval_d_lagrange(loc[1], i, p, n);
  }
}

/* src/eles_hexas.cpp:1167-1205 / src/eles_quads.cpp:1107-1128 (VCJH branch) */
static double div_vcjh_basis(const or_solver* s, int idx, const double* loc) {
  int P = s->order, n = P + 1;
  const double* p = s->loc1d;
  double eta = s->deck.vcjh_scheme == 0 ? s->deck.eta : compute_eta(s->deck.vcjh_scheme, P);
  if (s->ele_type == 4) {
    int i = idx / s->nfpf, j = (idx - s->nfpf * i) / n, k = idx - s->nfpf * i - n * j;
    switch (i) {
      case 0: return -eval_lagrange(loc[0], P - k, p, n) * eval_lagrange(loc[1], j, p, n) * eval_d_vcjh_1d(loc[2], 0, P, eta);
      case 1: return -eval_lagrange(loc[0], k, p, n) * eval_lagrange(loc[2], j, p, n) * eval_d_vcjh_1d(loc[1], 0, P, eta);
      case 2: return eval_lagrange(loc[1], k, p, n) * eval_lagrange(loc[2], j, p, n) * eval_d_vcjh_1d(loc[0], 1, P, eta);
      case 3: return eval_lagrange(loc[0], P - k, p, n) * eval_lagrange(loc[2], j, p, n) * eval_d_vcjh_1d(loc[1], 1, P, eta);
      case 4: return -eval_lagrange(loc[1], P - k, p, n) * eval_lagrange(loc[2], j, p, n) * eval_d_vcjh_1d(loc[0], 0, P, eta);
      default: return eval_lagrange(loc[0], k, p, n) * eval_lagrange(loc[1], j, p, n) * eval_d_vcjh_1d(loc[2], 1, P, eta);
    }
  } else {
    int i = idx / s->nfpf, j = idx - s->nfpf * i;
    switch (i) {
      case 0: return -eval_lagrange(loc[0], j, p, n) * eval_d_vcjh_1d(loc[1], 0, P, eta);
      case 1: return eval_lagrange(loc[1], j, p, n) * eval_d_vcjh_1d(loc[0], 1, P, eta);
      case 2: return eval_lagrange(loc[0], P - j, p, n) * eval_d_vcjh_1d(loc[1], 1, P, eta);
      default: return -eval_lagrange(loc[1], P - j, p, n) * eval_d_vcjh_1d(loc[0], 0, P, eta);
    }
  }
}

/* src/eles_hexas.cpp:240-256, 266-324, 572-625; src/eles_quads.cpp:237-250, 259-300, 443-480 */
static void set_points(or_solver* s) {
  int P = s->order, n = P + 1, d = s->nd;
  const double* g = s->loc1d;
  s->loc_upts = xcalloc((size_t)d * s->nu, sizeof(double));
  s->tloc_fpts = xcalloc((size_t)d * s->nfp, sizeof(double));
  s->tnorm_fpts = xcalloc((size_t)d * s->nfp, sizeof(double));
  if (s->ele_type == 2) {
    /* src/eles_tets.cpp:233-331 (interior points), 589-620 (normals) */
    const double* up = tet_pts_tab[P];
    const double* tp = tri_pts_tab[P];
    for (int u = 0; u < s->nu; u++)
      for (int k = 0; k < 3; k++) s->loc_upts[k + 3 * u] = up[3 * u + k];
    int nf1 = s->nfpf;
    for (int j = 0; j < P + 1; j++)
      for (int i = 0; i < P + 1 - j; i++) {
        int itri = j * (P + 1) - (j - 1) * j / 2 + i;
        int ialpha = j * (P + 1) - (j - 1) * j / 2 + (P - j - i);
        double* t0 = s->tloc_fpts + 3 * itri;
        t0[0] = tp[2 * ialpha + 0];
        ialpha = itri;
        t0[1] = tp[2 * ialpha + 0];
        t0[2] = tp[2 * ialpha + 1];
        double* t1 = s->tloc_fpts + 3 * (nf1 + itri);
        t1[0] = -1; t1[1] = tp[2 * ialpha + 1]; t1[2] = tp[2 * ialpha + 0];
        double* t2 = s->tloc_fpts + 3 * (2 * nf1 + itri);
        t2[0] = tp[2 * ialpha + 0]; t2[1] = -1; t2[2] = tp[2 * ialpha + 1];
        double* t3 = s->tloc_fpts + 3 * (3 * nf1 + itri);
        t3[0] = tp[2 * ialpha + 1]; t3[1] = tp[2 * ialpha + 0]; t3[2] = -1;
      }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < nf1; j++) {
        double* tn = s->tnorm_fpts + 3 * (nf1 * i + j);
        tn[0] = tn[1] = tn[2] = 0.0;
        if (i == 0) tn[0] = tn[1] = tn[2] = 1. / sqrt(3.);
        else tn[i - 1] = -1.0;
      }
    /* Vandermonde of the orthonormal basis and its inverse (set_vandermonde,
       src/eles_tets.cpp:722-733) */
    s->inv_vdm = xcalloc((size_t)s->nu * s->nu, sizeof(double));
    for (int i = 0; i < s->nu; i++)
      for (int j = 0; j < s->nu; j++)
        s->inv_vdm[i + s->nu * j] = dubiner_3d(s->loc_upts[3 * i], s->loc_upts[3 * i + 1], s->loc_upts[3 * i + 2], j);
    invert(s->inv_vdm, s->nu);
  } else if (s->ele_type == 4) {
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          int u = idx3(n, k, j, i);
          s->loc_upts[0 + 3 * u] = g[k]; s->loc_upts[1 + 3 * u] = g[j]; s->loc_upts[2 + 3 * u] = g[i];
        }
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          int f = k + n * j + n * n * i;
          double* t = s->tloc_fpts + 3 * f;
          double* tn = s->tnorm_fpts + 3 * f;
          tn[0] = tn[1] = tn[2] = 0.0;
          switch (i) {
            case 0: t[0] = g[P - k]; t[1] = g[j]; t[2] = -1.0; tn[2] = -1.0; break;
            case 1: t[0] = g[k]; t[1] = -1.0; t[2] = g[j]; tn[1] = -1.0; break;
            case 2: t[0] = 1.0; t[1] = g[k]; t[2] = g[j]; tn[0] = 1.0; break;
            case 3: t[0] = g[P - k]; t[1] = 1.0; t[2] = g[j]; tn[1] = 1.0; break;
            case 4: t[0] = -1.0; t[1] = g[P - k]; t[2] = g[j]; tn[0] = -1.0; break;
            default: t[0] = g[k]; t[1] = g[j]; t[2] = 1.0; tn[2] = 1.0; break;
          }
        }
  } else {
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        int u = j + n * i;
        s->loc_upts[0 + 2 * u] = g[j]; s->loc_upts[1 + 2 * u] = g[i];
      }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < n; j++) {
        int f = j + n * i;
        double* t = s->tloc_fpts + 2 * f;
        double* tn = s->tnorm_fpts + 2 * f;
        tn[0] = tn[1] = 0.0;
        switch (i) {
          case 0: t[0] = g[j]; t[1] = -1.0; tn[1] = -1.0; break;
          case 1: t[0] = 1.0; t[1] = g[j]; tn[0] = 1.0; break;
          case 2: t[0] = g[P - j]; t[1] = 1.0; tn[1] = 1.0; break;
          default: t[0] = -1.0; t[1] = g[P - j]; tn[0] = -1.0; break;
        }
      }
  }
}

/* src/eles.cpp:3620-4052 (set_opp_0 .. set_opp_6, dense forms) */
static void set_operators(or_solver* s) {
  int nu = s->nu, nfp = s->nfp, d = s->nd;
  s->opp_0 = xcalloc((size_t)nfp * nu, sizeof(double));
  s->opp_3 = xcalloc((size_t)nu * nfp, sizeof(double));
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < nfp; j++)
      s->opp_0[j + nfp * i] = nodal_basis(s, i, s->tloc_fpts + d * j);
  for (int dim = 0; dim < d; dim++) {
    s->opp_1[dim] = xcalloc((size_t)nfp * nu, sizeof(double));
    s->opp_2[dim] = xcalloc((size_t)nu * nu, sizeof(double));
    s->opp_4[dim] = xcalloc((size_t)nu * nu, sizeof(double));
    s->opp_5[dim] = xcalloc((size_t)nu * nfp, sizeof(double));
    for (int j = 0; j < nu; j++)
      for (int k = 0; k < nfp; k++)
        s->opp_1[dim][k + nfp * j] = nodal_basis(s, j, s->tloc_fpts + d * k) * s->tnorm_fpts[dim + d * k];
    for (int j = 0; j < nu; j++)
      for (int k = 0; k < nu; k++) {
        double v = d_nodal_basis(s, j, dim, s->loc_upts + d * k);
        s->opp_2[dim][k + nu * j] = v;
        s->opp_4[dim][k + nu * j] = v;
      }
  }
  /* fill_opp_3: src/eles_hexas.cpp:1148-1162; tets: Filt * opp_3_dg with the DG
     filter Filt = I (c_tet = 0), src/eles_tets.cpp:972-1016 */
  if (s->ele_type == 2) {
    tet_dg_lift(s, s->opp_3);
  } else {
    for (int i = 0; i < nfp; i++)
      for (int j = 0; j < nu; j++)
        s->opp_3[j + nu * i] = div_vcjh_basis(s, i, s->loc_upts + d * j);
  }
  for (int dim = 0; dim < d; dim++)
    for (int j = 0; j < nfp; j++)
      for (int k = 0; k < nu; k++)
        s->opp_5[dim][k + nu * j] = s->opp_3[k + nu * j] * s->tnorm_fpts[dim + d * j];
  s->opp_6 = s->opp_0;
}

/* Linear shape basis and its derivative, src/eles_hexas.cpp:1000-1020, 1075-1093;
 * src/eles_quads.cpp:960-975, 1000-1015 (perfect square/cube node counts) */
static void shape_basis(const or_solver* s, const double* loc, double* N, double* dN /* (nsp,d) */) {
  double sp[2];
  for (int i = 0; i < 2; i++) sp[i] = -1.0 + ((2.0 * i) / (1.0 * (2 - 1)));
  if (s->ele_type == 2) {  /* linear tet, src/eles_tets.cpp:862-880, 909-935 */
    if (N) {
      N[0] = -0.5 * (loc[0] + loc[1] + loc[2] + 1.);
      N[1] = 0.5 * (loc[0] + 1.);
      N[2] = 0.5 * (loc[1] + 1.);
      N[3] = 0.5 * (loc[2] + 1.);
    }
    if (dN) {
      for (int m = 0; m < 12; m++) dN[m] = 0.0;
      for (int k = 0; k < 3; k++) { dN[0 + 4 * k] = -0.5; dN[k + 1 + 4 * k] = 0.5; }
    }
    return;
  }
  if (s->ele_type == 4) {
    for (int m = 0; m < 8; m++) {
      int i = m / 4, j = (m - 4 * i) / 2, k = m - 2 * j - 4 * i;
      if (N) N[m] = eval_lagrange(loc[0], k, sp, 2) * eval_lagrange(loc[1], j, sp, 2) * eval_lagrange(loc[2], i, sp, 2);
      if (dN) {
        dN[m + 8 * 0] = eval_d_lagrange(loc[0], k, sp, 2) * eval_lagrange(loc[1], j, sp, 2) * eval_lagrange(loc[2], i, sp, 2);
        dN[m + 8 * 1] = eval_lagrange(loc[0], k, sp, 2) * eval_d_lagrange(loc[1], j, sp, 2) * eval_lagrange(loc[2], i, sp, 2);
        dN[m + 8 * 2] = eval_lagrange(loc[0], k, sp, 2) * eval_lagrange(loc[1], j, sp, 2) * eval_d_lagrange(loc[2], i, sp, 2);
      }
    }
  } else {
    for (int m = 0; m < 4; m++) {
      int i = m / 2, j = m - 2 * i;
      if (N) N[m] = eval_lagrange(loc[0], j, sp, 2) * eval_lagrange(loc[1], i, sp, 2);
      if (dN) {
        dN[m + 4 * 0] = eval_d_lagrange(loc[0], j, sp, 2) * eval_lagrange(loc[1], i, sp, 2);
        dN[m + 4 * 1] = eval_lagrange(loc[0], j, sp, 2) * eval_d_lagrange(loc[1], i, sp, 2);
      }
    }
  }
}

/* src/eles.cpp:5744-5757 calc_pos */
static void calc_pos(const or_solver* s, const double* loc, int e, double* pos) {
  double N[8];
  shape_basis(s, loc, N, NULL);
  for (int i = 0; i < s->nd; i++) {
    pos[i] = 0.0;
    for (int j = 0; j < s->nsp; j++) pos[i] += N[j] * s->shape[i + s->nd * (j + s->nsp * e)];
  }
}

/* src/eles.cpp:5841-5856 calc_d_pos: dpos(j,k) = dx_j/dr_k */
static void calc_d_pos(const or_solver* s, const double* loc, int e, double dpos[3][3]) {
  double dN[24];
  shape_basis(s, loc, NULL, dN);
  for (int j = 0; j < s->nd; j++)
    for (int k = 0; k < s->nd; k++) {
      dpos[j][k] = 0.0;
      for (int i = 0; i < s->nsp; i++) dpos[j][k] += dN[i + s->nsp * k] * s->shape[j + s->nd * (i + s->nsp * e)];
    }
}

#define JG(l, m, j, i) s->JGinv_upts[(l) + nd * ((m) + nd * ((j) + (long)nu * (i)))]

/* src/eles.cpp:4557-4847 set_transforms (static mesh) */
static void set_transforms(or_solver* s) {
  int nd = s->nd, nu = s->nu, nfp = s->nfp, N = s->n_eles;
  s->detjac_upts = xcalloc((size_t)nu * N, sizeof(double));
  s->JGinv_upts = xcalloc((size_t)nd * nd * nu * N, sizeof(double));
  s->pos_upts = xcalloc((size_t)nu * N * nd, sizeof(double));
  s->tdA_fpts = xcalloc((size_t)nfp * N, sizeof(double));
  s->norm_fpts = xcalloc((size_t)nfp * N * nd, sizeof(double));
  s->pos_fpts = xcalloc((size_t)nfp * N * nd, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double d[3][3], pos[3];
    for (int j = 0; j < nu; j++) {
      const double* loc = s->loc_upts + nd * j;
      calc_pos(s, loc, i, pos);
      for (int k = 0; k < nd; k++) s->pos_upts[j + nu * (i + (long)N * k)] = pos[k];
      calc_d_pos(s, loc, i, d);
      if (nd == 2) {
        double xr = d[0][0], xs = d[0][1], yr = d[1][0], ys = d[1][1];
        s->detjac_upts[j + nu * i] = xr * ys - xs * yr;
        JG(0, 0, j, i) = ys; JG(0, 1, j, i) = -xs; JG(1, 0, j, i) = -yr; JG(1, 1, j, i) = xr;
      } else {
        double xr = d[0][0], xs = d[0][1], xt = d[0][2], yr = d[1][0], ys = d[1][1], yt = d[1][2],
               zr = d[2][0], zs = d[2][1], zt = d[2][2];
        s->detjac_upts[j + nu * i] = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
        JG(0, 0, j, i) = ys * zt - yt * zs; JG(0, 1, j, i) = xt * zs - xs * zt; JG(0, 2, j, i) = xs * yt - xt * ys;
        JG(1, 0, j, i) = yt * zr - yr * zt; JG(1, 1, j, i) = xr * zt - xt * zr; JG(1, 2, j, i) = xt * yr - xr * yt;
        JG(2, 0, j, i) = yr * zs - ys * zr; JG(2, 1, j, i) = xs * zr - xr * zs; JG(2, 2, j, i) = xr * ys - xs * yr;
      }
    }
    for (int j = 0; j < nfp; j++) {
      const double* loc = s->tloc_fpts + nd * j;
      const double* tn = s->tnorm_fpts + nd * j;
      calc_pos(s, loc, i, pos);
      for (int k = 0; k < nd; k++) s->pos_fpts[j + nfp * (i + (long)N * k)] = pos[k];
      calc_d_pos(s, loc, i, d);
      double t[3], tdA;
      if (nd == 2) {
        t[0] = (tn[0] * d[1][1]) - (tn[1] * d[1][0]);
        t[1] = -(tn[0] * d[0][1]) + (tn[1] * d[0][0]);
        tdA = sqrt(t[0] * t[0] + t[1] * t[1]);
      } else {
        t[0] = ((tn[0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1])) + (tn[1] * (d[1][2] * d[2][0] - d[1][0] * d[2][2])) +
                (tn[2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0])));
        t[1] = ((tn[0] * (d[0][2] * d[2][1] - d[0][1] * d[2][2])) + (tn[1] * (d[0][0] * d[2][2] - d[0][2] * d[2][0])) +
                (tn[2] * (d[0][1] * d[2][0] - d[0][0] * d[2][1])));
        t[2] = ((tn[0] * (d[0][1] * d[1][2] - d[0][2] * d[1][1])) + (tn[1] * (d[0][2] * d[1][0] - d[0][0] * d[1][2])) +
                (tn[2] * (d[0][0] * d[1][1] - d[0][1] * d[1][0])));
        tdA = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      }
      s->tdA_fpts[j + nfp * i] = tdA;
      for (int k = 0; k < nd; k++) s->norm_fpts[j + nfp * (i + (long)N * k)] = t[k] / tdA;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Connectivity: face matching by shared vertices (src/geometry.cpp:2412-2700),
 * periodic faces by translated centroids (src/geometry.cpp:489-563), flux points
 * paired by position (the reference's rot_tag LUT, src/inters.cpp:187-296, encodes
 * the same pairing). */

static const int hex_face_nodes[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {1, 3, 5, 7},
                                         {2, 3, 6, 7}, {0, 2, 4, 6}, {4, 5, 6, 7}};
static const int quad_face_nodes[4][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}};
/* src/geometry.cpp:3156-3185 (linear tet) */
static const int tet_face_nodes[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

static int face_node(const or_solver* s, int f, int k) {
  if (s->ele_type == 4) return hex_face_nodes[f][k];
  if (s->ele_type == 2) return tet_face_nodes[f][k];
  return quad_face_nodes[f][k];
}
static int face_nv(const or_solver* s) { return s->ele_type == 4 ? 4 : (s->ele_type == 2 ? 3 : 2); }

typedef struct { int key[4]; int e, f; } face_rec;

static int cmp_face(const void* a, const void* b) {
  const face_rec *x = a, *y = b;
  for (int i = 0; i < 4; i++) {
    if (x->key[i] != y->key[i]) return x->key[i] < y->key[i] ? -1 : 1;
  }
  if (x->e != y->e) return x->e < y->e ? -1 : 1;
  return x->f - y->f;
}

static void sort4(int* k, int n) {
  for (int i = 1; i < n; i++)
    for (int j = i; j > 0 && k[j - 1] > k[j]; j--) { int t = k[j]; k[j] = k[j - 1]; k[j - 1] = t; }
}

static void face_centroid(const or_solver* s, const int* c2v, const double* xv, int e, int f, double* c) {
  int nv = face_nv(s);
  int nsp = s->nsp;
  for (int m = 0; m < s->nd; m++) c[m] = 0.0;
  for (int k = 0; k < nv; k++) {
    int node = face_node(s, f, k);
    int v = c2v[node + nsp * e];
    for (int m = 0; m < s->nd; m++) c[m] += xv[m + s->nd * v] / nv;
  }
}

/* Pair the flux points of (eL,fL) with those of (eR,fR); shift = xR - xL. */
static void pair_points(or_solver* s, int eL, int fL, int eR, int fR, const double* shift, int* outL, int* outR) {
  int nfpf = s->nfpf, nfp = s->nfp, N = s->n_eles, nd = s->nd;
  for (int j = 0; j < nfpf; j++) {
    int fl = fL * nfpf + j;
    double pl[3];
    for (int m = 0; m < nd; m++) pl[m] = s->pos_fpts[fl + nfp * (eL + (long)N * m)] + shift[m];
    int best = -1;
    double bd = 1e300;
    for (int q = 0; q < nfpf; q++) {
      int fr = fR * nfpf + q;
      double dd = 0;
      for (int m = 0; m < nd; m++) {
        double t = s->pos_fpts[fr + nfp * (eR + (long)N * m)] - pl[m];
        dd += t * t;
      }
      if (dd < bd) { bd = dd; best = fr; }
    }
    outL[j] = fl + nfp * eL;
    outR[j] = best + nfp * eR;
  }
}

/* is_bdy: (N*nfaces) 1 for faces on a physical boundary (not matched) */
static int set_connectivity(or_solver* s, const int* c2v, const double* xv, const char* is_bdy) {
  int N = s->n_eles, nfaces = s->nfaces, nv = face_nv(s);
  long nrec = 0;
  for (long i = 0; i < (long)N * nfaces; i++) nrec += !is_bdy[i];
  face_rec* r = xcalloc(nrec, sizeof(face_rec));
  long ir = 0;
  for (int e = 0; e < N; e++)
    for (int f = 0; f < nfaces; f++) {
      if (is_bdy[(long)e * nfaces + f]) continue;
      face_rec* q = r + ir++;
      q->e = e; q->f = f;
      for (int k = 0; k < 4; k++) q->key[k] = -1;
      for (int k = 0; k < nv; k++) {
        int node = face_node(s, f, k);
        q->key[k] = c2v[node + s->nsp * e];
      }
      sort4(q->key, nv);
    }
  qsort(r, nrec, sizeof(face_rec), cmp_face);
  /* partner[e*nfaces+f] = other (e,f) or -1; boundary faces -2 */
  long nall = (long)N * nfaces;
  long* partner = xcalloc(nall, sizeof(long));
  for (long i = 0; i < nall; i++) partner[i] = is_bdy[i] ? -2 : -1;
  for (long i = 0; i + 1 < nrec; i++) {
    if (memcmp(r[i].key, r[i + 1].key, sizeof(r[i].key)) == 0) {
      long a = (long)r[i].e * nfaces + r[i].f, b = (long)r[i + 1].e * nfaces + r[i + 1].f;
      partner[a] = b; partner[b] = a;
      i++;
    }
  }
  free(r);
  /* periodic matching of the remaining faces */
  double delta[3] = {s->deck.dx_cyclic, s->deck.dy_cyclic, s->deck.dz_cyclic};
  long nun = 0;
  long* un = xcalloc(nall, sizeof(long));
  for (long i = 0; i < nall; i++) if (partner[i] == -1) un[nun++] = i;
  double* cen = xcalloc((size_t)nun * 3, sizeof(double));
  for (long i = 0; i < nun; i++) face_centroid(s, c2v, xv, (int)(un[i] / nfaces), (int)(un[i] % nfaces), cen + 3 * i);
  const double tol = 1.e-6;
  for (long i = 0; i < nun; i++) {
    if (partner[un[i]] >= 0) continue;
    for (long j = i + 1; j < nun; j++) {
      if (partner[un[j]] >= 0) continue;
      int ok = 0;
      for (int m = 0; m < s->nd && !ok; m++) {
        if (!isfinite(delta[m])) continue;
        int match = fabs(fabs(cen[3 * i + m] - cen[3 * j + m]) - delta[m]) < tol;
        for (int q = 0; q < s->nd && match; q++)
          if (q != m && fabs(cen[3 * i + q] - cen[3 * j + q]) >= tol) match = 0;
        ok = match;
      }
      if (ok) { partner[un[i]] = un[j]; partner[un[j]] = un[i]; break; }
    }
  }
  for (long i = 0; i < nall; i++)
    if (partner[i] == -1) {
      fprintf(stderr, "fr_oracle: unmatched face (ele %ld, face %ld) without boundary condition\n",
              i / nfaces, i % nfaces);
      free(partner); free(un); free(cen);
      return -1;
    }
  /* interfaces in order of first appearance: L = first (e,f) in element order */
  s->n_inters = (int)(nrec / 2);
  s->face_l = xcalloc((size_t)s->n_inters * s->nfpf, sizeof(int));
  s->face_r = xcalloc((size_t)s->n_inters * s->nfpf, sizeof(int));
  int it = 0;
  for (long a = 0; a < nall; a++) {
    long b = partner[a];
    if (b < a) continue;
    int eL = (int)(a / nfaces), fL = (int)(a % nfaces), eR = (int)(b / nfaces), fR = (int)(b % nfaces);
    double cl[3] = {0, 0, 0}, cr[3] = {0, 0, 0}, shift[3] = {0, 0, 0};
    face_centroid(s, c2v, xv, eL, fL, cl);
    face_centroid(s, c2v, xv, eR, fR, cr);
    for (int m = 0; m < s->nd; m++) {
      double dlt = cr[m] - cl[m];
      if (isfinite(delta[m]) && fabs(fabs(dlt) - delta[m]) < tol) shift[m] = dlt > 0 ? delta[m] : -delta[m];
    }
    pair_points(s, eL, fL, eR, fR, shift, s->face_l + (long)it * s->nfpf, s->face_r + (long)it * s->nfpf);
    it++;
  }
  free(partner); free(un); free(cen);
  return 0;
}

/* ------------------------------------------------------------------------- */

or_solver* or_create(const or_deck* deck, int ele_type, int n_verts, const double* xv, int n_eles,
                     const int* c2v) {
  return or_create_bc(deck, ele_type, n_verts, xv, n_eles, c2v, 0, NULL, NULL, NULL);
}

/* Boundary faces (element, local face in the reference's numbering, bctype) are
 * taken out of the face matching and become boundary interfaces
 * (bdy_inters::set_boundary, src/bdy_inters.cpp:114-148); bctype 11 only. */
or_solver* or_create_bc(const or_deck* deck, int ele_type, int n_verts, const double* xv, int n_eles,
                        const int* c2v, int n_bfaces, const int* bf_ele, const int* bf_face,
                        const int* bf_type) {
  if (ele_type != 1 && ele_type != 2 && ele_type != 4) { fprintf(stderr, "fr_oracle: ele_type %d\n", ele_type); return NULL; }
  if (deck->order < 0 || deck->order > 6 || deck->upts_type != 0 || deck->equation != 0 ||
      (deck->LES && (!deck->viscous || !(deck->SGS_model == 0 || deck->SGS_model == 1 ||
                                         deck->SGS_model == 10))) ||
      (ele_type == 2 && (deck->order < 1 || deck->order > 5 || !(deck->vcjh_scheme == 1 ||
                                                                 (deck->vcjh_scheme == 0 && deck->eta == 0.0))))) {
    /* tets: orders 1-5 of the tabulated point sets, DG correction (c_tet = 0) */
    fprintf(stderr, "fr_oracle: unsupported deck\n");
    return NULL;
  }
  or_solver* s = xcalloc(1, sizeof(or_solver));
  s->deck = *deck;
  or_setup_params(deck, &s->ph);
  s->ele_type = ele_type;
  s->order = deck->order;
  int n = s->order + 1;
  s->nd = ele_type == 1 ? 2 : 3;
  s->nf = s->nd + 2;
  if (ele_type == 2) {  /* src/eles_tets.cpp:68-141 */
    s->nu = n * (n + 1) * (n + 2) / 6;
    s->nfaces = 4;
    s->nfpf = n * (n + 1) / 2;
    s->nsp = 4;
  } else {
    s->nu = ele_type == 4 ? n * n * n : n * n;
    s->nfaces = ele_type == 4 ? 6 : 4;
    s->nfpf = ele_type == 4 ? n * n : n;
    s->nsp = ele_type == 4 ? 8 : 4;
  }
  s->nfp = s->nfaces * s->nfpf;
  s->n_eles = n_eles;
  s->n_verts = n_verts;
  for (int i = 0; i < n; i++) s->loc1d[i] = gauss_tab[s->order][i];
  set_points(s);
  set_operators(s);
  s->shape = xcalloc((size_t)s->nd * s->nsp * n_eles, sizeof(double));
  for (int e = 0; e < n_eles; e++)
    for (int j = 0; j < s->nsp; j++)
      for (int m = 0; m < s->nd; m++)
        s->shape[m + s->nd * (j + s->nsp * e)] = xv[m + s->nd * c2v[j + s->nsp * e]];
  set_transforms(s);
  char* is_bdy = xcalloc((size_t)n_eles * s->nfaces, 1);
  for (int q = 0; q < n_bfaces; q++) {
    if (bf_ele[q] < 0 || bf_ele[q] >= n_eles || bf_face[q] < 0 || bf_face[q] >= s->nfaces || bf_type[q] != 11 ||
        !deck->viscous) {
      fprintf(stderr, "fr_oracle: boundary face %d not supported (isothermal walls, viscous only)\n", q);
      free(is_bdy); or_destroy(s); return NULL;
    }
    is_bdy[(long)bf_ele[q] * s->nfaces + bf_face[q]] = 1;
  }
  s->n_bdy = n_bfaces * s->nfpf;
  s->bdy_fpt = xcalloc((size_t)s->n_bdy, sizeof(int));
  s->bdy_type = xcalloc((size_t)s->n_bdy, sizeof(int));
  for (int q = 0; q < n_bfaces; q++)
    for (int j = 0; j < s->nfpf; j++) {
      s->bdy_fpt[q * s->nfpf + j] = bf_face[q] * s->nfpf + j + s->nfp * bf_ele[q];
      s->bdy_type[q * s->nfpf + j] = bf_type[q];
    }
  int cerr = set_connectivity(s, c2v, xv, is_bdy);
  free(is_bdy);
  if (cerr != 0) { or_destroy(s); return NULL; }
  long nuN = (long)s->nu * n_eles, nfpN = (long)s->nfp * n_eles;
  int nf = s->nf, nd = s->nd;
  s->disu_upts[0] = xcalloc(nuN * nf, sizeof(double));
  s->disu_upts[1] = xcalloc(nuN * nf, sizeof(double));
  s->div_tconf_upts = xcalloc(nuN * nf, sizeof(double));
  s->src_upts = xcalloc(nuN * nf, sizeof(double));
  s->disu_fpts = xcalloc(nfpN * nf, sizeof(double));
  s->tdisf_upts = xcalloc(nuN * nf * nd, sizeof(double));
  s->norm_tdisf_fpts = xcalloc(nfpN * nf, sizeof(double));
  s->norm_tconf_fpts = xcalloc(nfpN * nf, sizeof(double));
  if (deck->viscous) {
    s->delta_disu_fpts = xcalloc(nfpN * nf, sizeof(double));
    s->grad_disu_upts = xcalloc(nuN * nf * nd, sizeof(double));
    s->grad_disu_fpts = xcalloc(nfpN * nf * nd, sizeof(double));
  }
  if (deck->LES) {
    s->sgsf_upts = xcalloc(nuN * nf * nd, sizeof(double));
    s->sgsf_fpts = xcalloc(nfpN * nf * nd, sizeof(double));
    s->sgs_nu = xcalloc(nuN, sizeof(double));
  }
  return s;
}

void or_destroy(or_solver* s) {
  if (!s) return;
  free(s->loc_upts); free(s->tloc_fpts); free(s->tnorm_fpts); free(s->opp_0); free(s->opp_3);
  for (int d = 0; d < 3; d++) { free(s->opp_1[d]); free(s->opp_2[d]); free(s->opp_4[d]); free(s->opp_5[d]); }
  free(s->shape); free(s->detjac_upts); free(s->JGinv_upts); free(s->tdA_fpts); free(s->norm_fpts);
  free(s->pos_upts); free(s->pos_fpts);
  free(s->disu_upts[0]); free(s->disu_upts[1]); free(s->div_tconf_upts); free(s->src_upts);
  free(s->disu_fpts); free(s->tdisf_upts); free(s->norm_tdisf_fpts); free(s->norm_tconf_fpts);
  free(s->delta_disu_fpts); free(s->grad_disu_upts); free(s->grad_disu_fpts);
  free(s->sgsf_upts); free(s->sgsf_fpts); free(s->sgs_nu);
  free(s->face_l); free(s->face_r); free(s->bdy_fpt); free(s->bdy_type); free(s->inv_vdm);
  free(s);
}

int or_n_upts(const or_solver* s) { return s->nu; }
int or_n_fpts(const or_solver* s) { return s->nfp; }
int or_n_fields(const or_solver* s) { return s->nf; }
int or_n_dims(const or_solver* s) { return s->nd; }
int or_n_eles(const or_solver* s) { return s->n_eles; }
int or_n_inters(const or_solver* s) { return s->n_inters; }
void or_get_phys(const or_solver* s, or_phys* out) { *out = s->ph; }

long or_get(const or_solver* s, int what, double* out) {
  long nuN = (long)s->nu * s->n_eles, nfpN = (long)s->nfp * s->n_eles;
  int nf = s->nf, nd = s->nd;
  const double* src = NULL;
  long n = 0;
  switch (what) {
    case 0: src = s->disu_upts[0]; n = nuN * nf; break;
    case 1: src = s->disu_upts[1]; n = nuN * nf; break;
    case 2: src = s->div_tconf_upts; n = nuN * nf; break;
    case 3: src = s->disu_fpts; n = nfpN * nf; break;
    case 4: src = s->grad_disu_upts; n = src ? nuN * nf * nd : 0; break;
    case 5: src = s->grad_disu_fpts; n = src ? nfpN * nf * nd : 0; break;
    case 6: src = s->detjac_upts; n = nuN; break;
    case 7: src = s->JGinv_upts; n = nuN * nd * nd; break;
    case 8: src = s->tdA_fpts; n = nfpN; break;
    case 9: src = s->norm_fpts; n = nfpN * nd; break;
    case 10: src = s->pos_upts; n = nuN * nd; break;
    case 11: src = s->norm_tconf_fpts; n = nfpN * nf; break;
    case 12: src = s->opp_0; n = (long)s->nfp * s->nu; break;
    case 13: src = s->opp_3; n = (long)s->nu * s->nfp; break;
    case 14: n = (long)s->nu * s->nu * nd; break;
    case 15: src = s->loc_upts; n = (long)nd * s->nu; break;
    case 16: src = s->tloc_fpts; n = (long)nd * s->nfp; break;
    case 17: src = s->pos_fpts; n = nfpN * nd; break;
    case 18: src = s->sgsf_upts; n = src ? nuN * nf * nd : 0; break;
    case 19: src = s->sgsf_fpts; n = src ? nfpN * nf * nd : 0; break;
    default: return -1;
  }
  if (out && n > 0) {
    if (what == 14) {
      for (int d = 0; d < nd; d++) memcpy(out + (long)d * s->nu * s->nu, s->opp_4[d], sizeof(double) * s->nu * s->nu);
    } else {
      memcpy(out, src, sizeof(double) * n);
    }
  }
  return n;
}

void or_set_state(or_solver* s, const double* u) {
  memcpy(s->disu_upts[0], u, sizeof(double) * (long)s->nu * s->n_eles * s->nf);
}
void or_zero_rk_register(or_solver* s) {
  memset(s->disu_upts[1], 0, sizeof(double) * (long)s->nu * s->n_eles * s->nf);
}

long or_get_face_pairs(const or_solver* s, int* left, int* right) {
  long n = (long)s->n_inters * s->nfpf;
  if (left) memcpy(left, s->face_l, sizeof(int) * n);
  if (right) memcpy(right, s->face_r, sizeof(int) * n);
  return n;
}

/* ------------------------------------------------------------------------- */
/* Pointwise physics, src/flux.cpp */

/* src/flux.cpp:36-75 (2-D) and 77-128 (3-D), Euler part; f is (nf,nd) col-major */
static void calc_invf(int nd, double gamma, const double* u, double* f) {
  int nf = nd + 2;
#define F(k, l) f[(k) + nf * (l)]
  if (nd == 2) {
    double vx = u[1] / u[0], vy = u[2] / u[0];
    double p = (gamma - 1.0) * (u[3] - (0.5 * u[0] * ((vx * vx) + (vy * vy))));
    F(0, 0) = u[1]; F(1, 0) = p + (u[1] * vx); F(2, 0) = u[2] * vx; F(3, 0) = vx * (u[3] + p);
    F(0, 1) = u[2]; F(1, 1) = u[1] * vy; F(2, 1) = p + (u[2] * vy); F(3, 1) = vy * (u[3] + p);
  } else {
    double vx = u[1] / u[0], vy = u[2] / u[0], vz = u[3] / u[0];
    double p = (gamma - 1.0) * (u[4] - (0.5 * u[0] * ((vx * vx) + (vy * vy) + (vz * vz))));
    F(0, 0) = u[1]; F(1, 0) = p + (u[1] * vx); F(2, 0) = u[2] * vx; F(3, 0) = u[3] * vx; F(4, 0) = vx * (u[4] + p);
    F(0, 1) = u[2]; F(1, 1) = u[1] * vy; F(2, 1) = p + (u[2] * vy); F(3, 1) = u[3] * vy; F(4, 1) = vy * (u[4] + p);
    F(0, 2) = u[3]; F(1, 2) = u[1] * vz; F(2, 2) = u[2] * vz; F(3, 2) = p + (u[3] * vz); F(4, 2) = vz * (u[4] + p);
  }
#undef F
}

/* src/flux.cpp:132-259 (2-D) and 263-431 (3-D), NS viscous flux without SA.
 * g is (nf,nd) col-major: g(k,m) = d u_k / d x_m */
static void calc_visf(int nd, const or_phys* ph, const double* u, const double* g, double* f) {
  int nf = nd + 2;
#define F(k, l) f[(k) + nf * (l)]
#define G(k, l) g[(k) + nf * (l)]
  double mu_t = 0.0;
  if (nd == 2) {
    double rho = u[0], mom_x = u[1], mom_y = u[2], ene = u[3];
    double rho_dx = G(0, 0), mom_x_dx = G(1, 0), mom_y_dx = G(2, 0), ene_dx = G(3, 0);
    double rho_dy = G(0, 1), mom_x_dy = G(1, 1), mom_y_dy = G(2, 1), ene_dy = G(3, 1);
    double uu = mom_x / rho, v = mom_y / rho;
    double inte = ene / rho - 0.5 * (uu * uu + v * v);
    double rt_ratio = (ph->gamma - 1.0) * inte / (ph->rt_inf);
    double mu = (ph->mu_inf) * pow(rt_ratio, 1.5) * (1. + (ph->c_sth)) / (rt_ratio + (ph->c_sth));
    mu = mu + ph->fix_vis * (ph->mu_inf - mu);
    double du_dx = (mom_x_dx - rho_dx * uu) / rho, du_dy = (mom_x_dy - rho_dy * uu) / rho;
    double dv_dx = (mom_y_dx - rho_dx * v) / rho, dv_dy = (mom_y_dy - rho_dy * v) / rho;
    double dke_dx = 0.5 * (uu * uu + v * v) * rho_dx + rho * (uu * du_dx + v * dv_dx);
    double dke_dy = 0.5 * (uu * uu + v * v) * rho_dy + rho * (uu * du_dy + v * dv_dy);
    double de_dx = (ene_dx - dke_dx - rho_dx * inte) / rho;
    double de_dy = (ene_dy - dke_dy - rho_dy * inte) / rho;
    double diag = (du_dx + dv_dy) / 3.0;
    double tauxx = 2.0 * (mu + mu_t) * (du_dx - diag);
    double tauxy = (mu + mu_t) * (du_dy + dv_dx);
    double tauyy = 2.0 * (mu + mu_t) * (dv_dy - diag);
    double kq = (mu / ph->prandtl + (mu_t == 0.0 ? 0.0 : mu_t / ph->prandtl_t)) * (ph->gamma);
    F(0, 0) = 0.0; F(1, 0) = -tauxx; F(2, 0) = -tauxy; F(3, 0) = -(uu * tauxx + v * tauxy + kq * de_dx);
    F(0, 1) = 0.0; F(1, 1) = -tauxy; F(2, 1) = -tauyy; F(3, 1) = -(uu * tauxy + v * tauyy + kq * de_dy);
  } else {
    double rho = u[0], mom_x = u[1], mom_y = u[2], mom_z = u[3], ene = u[4];
    double rho_dx = G(0, 0), mom_x_dx = G(1, 0), mom_y_dx = G(2, 0), mom_z_dx = G(3, 0), ene_dx = G(4, 0);
    double rho_dy = G(0, 1), mom_x_dy = G(1, 1), mom_y_dy = G(2, 1), mom_z_dy = G(3, 1), ene_dy = G(4, 1);
    double rho_dz = G(0, 2), mom_x_dz = G(1, 2), mom_y_dz = G(2, 2), mom_z_dz = G(3, 2), ene_dz = G(4, 2);
    double uu = mom_x / rho, v = mom_y / rho, w = mom_z / rho;
    double inte = ene / rho - 0.5 * (uu * uu + v * v + w * w);
    double rt_ratio = (ph->gamma - 1.0) * inte / (ph->rt_inf);
    double mu = (ph->mu_inf) * pow(rt_ratio, 1.5) * (1 + (ph->c_sth)) / (rt_ratio + (ph->c_sth));
    mu = mu + ph->fix_vis * (ph->mu_inf - mu);
    double du_dx = (mom_x_dx - rho_dx * uu) / rho, du_dy = (mom_x_dy - rho_dy * uu) / rho, du_dz = (mom_x_dz - rho_dz * uu) / rho;
    double dv_dx = (mom_y_dx - rho_dx * v) / rho, dv_dy = (mom_y_dy - rho_dy * v) / rho, dv_dz = (mom_y_dz - rho_dz * v) / rho;
    double dw_dx = (mom_z_dx - rho_dx * w) / rho, dw_dy = (mom_z_dy - rho_dy * w) / rho, dw_dz = (mom_z_dz - rho_dz * w) / rho;
    double dke_dx = 0.5 * (uu * uu + v * v + w * w) * rho_dx + rho * (uu * du_dx + v * dv_dx + w * dw_dx);
    double dke_dy = 0.5 * (uu * uu + v * v + w * w) * rho_dy + rho * (uu * du_dy + v * dv_dy + w * dw_dy);
    double dke_dz = 0.5 * (uu * uu + v * v + w * w) * rho_dz + rho * (uu * du_dz + v * dv_dz + w * dw_dz);
    double de_dx = (ene_dx - dke_dx - rho_dx * inte) / rho;
    double de_dy = (ene_dy - dke_dy - rho_dy * inte) / rho;
    double de_dz = (ene_dz - dke_dz - rho_dz * inte) / rho;
    double diag = (du_dx + dv_dy + dw_dz) / 3.0;
    double tauxx = 2.0 * (mu + mu_t) * (du_dx - diag);
    double tauyy = 2.0 * (mu + mu_t) * (dv_dy - diag);
    double tauzz = 2.0 * (mu + mu_t) * (dw_dz - diag);
    double tauxy = (mu + mu_t) * (du_dy + dv_dx);
    double tauxz = (mu + mu_t) * (du_dz + dw_dx);
    double tauyz = (mu + mu_t) * (dv_dz + dw_dy);
    double kq = (mu / ph->prandtl + (mu_t == 0.0 ? 0.0 : mu_t / ph->prandtl_t)) * (ph->gamma);
    F(0, 0) = 0.0; F(1, 0) = -tauxx; F(2, 0) = -tauxy; F(3, 0) = -tauxz;
    F(4, 0) = -(uu * tauxx + v * tauxy + w * tauxz + kq * de_dx);
    F(0, 1) = 0.0; F(1, 1) = -tauxy; F(2, 1) = -tauyy; F(3, 1) = -tauyz;
    F(4, 1) = -(uu * tauxy + v * tauyy + w * tauyz + kq * de_dy);
    F(0, 2) = 0.0; F(1, 2) = -tauxz; F(2, 2) = -tauyz; F(3, 2) = -tauzz;
    F(4, 2) = -(uu * tauxz + v * tauyz + w * tauzz + kq * de_dz);
  }
#undef F
#undef G
}

/* src/inters.cpp:311-365, with grid velocity zero (static mesh) */
static void rusanov_flux(int nd, double gamma, const double* ul, const double* ur, const double* fl,
                         const double* fr, const double* n, double* fn) {
  int nf = nd + 2;
  double fnl[5], fnr[5];
  for (int k = 0; k < nf; k++) {
    fnl[k] = 0.; fnr[k] = 0.;
    for (int l = 0; l < nd; l++) { fnl[k] += fl[k + nf * l] * n[l]; fnr[k] += fr[k + nf * l] * n[l]; }
  }
  double vx_l = ul[1] / ul[0], vx_r = ur[1] / ur[0], vy_l = ul[2] / ul[0], vy_r = ur[2] / ur[0];
  double vn_l, vn_r, vn_g, p_l, p_r;
  const double vg0 = 0.0;
  if (nd == 2) {
    vn_l = vx_l * n[0] + vy_l * n[1];
    vn_r = vx_r * n[0] + vy_r * n[1];
    vn_g = vg0 * n[0] + vg0 * n[1];
    p_l = (gamma - 1.0) * (ul[3] - (0.5 * ul[0] * ((vx_l * vx_l) + (vy_l * vy_l))));
    p_r = (gamma - 1.0) * (ur[3] - (0.5 * ur[0] * ((vx_r * vx_r) + (vy_r * vy_r))));
  } else {
    double vz_l = ul[3] / ul[0], vz_r = ur[3] / ur[0];
    vn_l = vx_l * n[0] + vy_l * n[1] + vz_l * n[2];
    vn_r = vx_r * n[0] + vy_r * n[1] + vz_r * n[2];
    vn_g = vg0 * n[0] + vg0 * n[1] + vg0 * n[2];
    p_l = (gamma - 1.0) * (ul[4] - (0.5 * ul[0] * ((vx_l * vx_l) + (vy_l * vy_l) + (vz_l * vz_l))));
    p_r = (gamma - 1.0) * (ur[4] - (0.5 * ur[0] * ((vx_r * vx_r) + (vy_r * vy_r) + (vz_r * vz_r))));
  }
  double vn_av_mag = sqrt(0.25 * (vn_l + vn_r) * (vn_l + vn_r));
  double c_av = sqrt((gamma * (p_l + p_r)) / (ul[0] + ur[0]));
  double eig = fabs(vn_av_mag - vn_g + c_av);
  for (int k = 0; k < nf; k++) fn[k] = 0.5 * ((fnl[k] + fnr[k]) - eig * (ur[k] - ul[k]));
}

/* Roe flux with Harten's entropy fix, src/inters.cpp:390-520, static mesh (grid
 * velocity zero).  The reference evaluates the first (central) part and the
 * pressure in 2-D only and stops in 3-D; its dissipation part is written for both.
 * Here the 3-D first part is the same expression with the third velocity
 * component: rho u.n, rho u.n v_i + p n_i, rho u.n h. */
static void roe_flux(int nd, double gamma, const double* ul, const double* ur, const double* n, double* fn) {
  int nf = nd + 2;
  double v_l[3] = {0, 0, 0}, v_r[3] = {0, 0, 0}, um[3] = {0, 0, 0}, du[5];
  for (int i = 0; i < nd; i++) { v_l[i] = ul[i + 1] / ul[0]; v_r[i] = ur[i + 1] / ur[0]; }
  double p_l, p_r;
  if (nd == 2) {
    p_l = (gamma - 1.0) * (ul[3] - (0.5 * ul[0] * ((v_l[0] * v_l[0]) + (v_l[1] * v_l[1]))));
    p_r = (gamma - 1.0) * (ur[3] - (0.5 * ur[0] * ((v_r[0] * v_r[0]) + (v_r[1] * v_r[1]))));
  } else {
    p_l = (gamma - 1.0) * (ul[4] - (0.5 * ul[0] * ((v_l[0] * v_l[0]) + (v_l[1] * v_l[1]) + (v_l[2] * v_l[2]))));
    p_r = (gamma - 1.0) * (ur[4] - (0.5 * ur[0] * ((v_r[0] * v_r[0]) + (v_r[1] * v_r[1]) + (v_r[2] * v_r[2]))));
  }
  double h_l = (ul[nd + 1] + p_l) / ul[0], h_r = (ur[nd + 1] + p_r) / ur[0];
  double sq_rho = sqrt(ur[0] / ul[0]);
  double rrho = 1. / (sq_rho + 1.);
  for (int i = 0; i < nd; i++) um[i] = rrho * (v_l[i] + sq_rho * v_r[i]);
  double hm = rrho * (h_l + sq_rho * h_r);
  double usq = 0.;
  for (int i = 0; i < nd; i++) usq += 0.5 * um[i] * um[i];
  double am_sq = (gamma - 1.) * (hm - usq);
  double am = sqrt(am_sq);
  double unm = 0., vgn = 0.;
  for (int i = 0; i < nd; i++) { unm += um[i] * n[i]; vgn += 0.0 * n[i]; }
  double rhoun_l = 0., rhoun_r = 0.;
  for (int i = 0; i < nd; i++) { rhoun_l += ul[i + 1] * n[i]; rhoun_r += ur[i + 1] * n[i]; }
  fn[0] = rhoun_l + rhoun_r;
  for (int i = 0; i < nd; i++) fn[i + 1] = rhoun_l * v_l[i] + rhoun_r * v_r[i] + (p_l + p_r) * n[i];
  fn[nd + 1] = rhoun_l * h_l + rhoun_r * h_r;
  for (int i = 0; i < nf; i++) du[i] = ur[i] - ul[i];
  double lambda0 = fabs(unm - vgn), lambdaP = fabs(unm - vgn + am), lambdaM = fabs(unm - vgn - am);
  double eps = 0.5 * (fabs(rhoun_l / ul[0] - rhoun_r / ur[0]) + fabs(sqrt(gamma * p_l / ul[0]) - sqrt(gamma * p_r / ur[0])));
  if (lambda0 < 2. * eps) lambda0 = 0.25 * lambda0 * lambda0 / eps + eps;
  if (lambdaP < 2. * eps) lambdaP = 0.25 * lambdaP * lambdaP / eps + eps;
  if (lambdaM < 2. * eps) lambdaM = 0.25 * lambdaM * lambdaM / eps + eps;
  double a2 = 0.5 * (lambdaP + lambdaM) - lambda0;
  double a3 = 0.5 * (lambdaP - lambdaM) / am;
  double a1 = a2 * (gamma - 1.) / am_sq;
  double a4 = a3 * (gamma - 1.);
  double a5, a6;
  if (nd == 2) {
    a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] + du[3];
    a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2];
  } else {
    a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] - um[2] * du[3] + du[4];
    a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2] - n[2] * du[3];
  }
  double aL1 = a1 * a5 - a3 * a6;
  double bL1 = a4 * a5 - a2 * a6;
  fn[0] = fn[0] - (lambda0 * du[0] + aL1);
  for (int i = 0; i < nd; i++) fn[i + 1] = fn[i + 1] - (lambda0 * du[i + 1] + aL1 * um[i] + bL1 * n[i]);
  fn[nd + 1] = fn[nd + 1] - (lambda0 * du[nd + 1] + aL1 * hm + bL1 * unm);
  for (int i = 0; i < nf; i++) fn[i] = 0.5 * fn[i] - 0.5 * vgn * (ur[i] + ul[i]);
}

void or_common_flux(int nd, int riemann, double gamma, const double* ul, const double* ur, const double* n,
                    double* fn) {
  if (riemann == 2) {
    roe_flux(nd, gamma, ul, ur, n, fn);
  } else {
    double fl[15], fr[15];
    calc_invf(nd, gamma, ul, fl);
    calc_invf(nd, gamma, ur, fr);
    rusanov_flux(nd, gamma, ul, ur, fl, fr, n, fn);
  }
}

/* LDG switch direction, src/inters.cpp:556-566 and 650-660 */
static double ldg_pen(int nd, double pen, const double* n) {
  if (nd == 2) { if ((n[0] + n[1]) < 0.) pen = -pen; }
  else { if ((n[0] + n[1] + sqrt(2.) * n[2]) < 0.) pen = -pen; }
  return pen;
}

/* src/inters.cpp:550-641 (flux_spec 0) */
static void ldg_flux(int nd, const double* ul, const double* ur, const double* fl, const double* fr,
                     const double* n, double* fn, double tau, double pen) {
  int nf = nd + 2;
  double fc[15];
  pen = ldg_pen(nd, pen, n);
  double nx = n[0], ny = n[1], nz = nd == 3 ? n[2] : 0.0;
#define FL(k, l) fl[(k) + nf * (l)]
#define FR(k, l) fr[(k) + nf * (l)]
  for (int k = 0; k < nf; k++) {
    if (nd == 2) {
      fc[k + nf * 0] = 0.5 * (FL(k, 0) + FR(k, 0)) + pen * nx * (nx * (FL(k, 0) - FR(k, 0)) + ny * (FL(k, 1) - FR(k, 1))) + tau * nx * (ul[k] - ur[k]);
      fc[k + nf * 1] = 0.5 * (FL(k, 1) + FR(k, 1)) + pen * ny * (nx * (FL(k, 0) - FR(k, 0)) + ny * (FL(k, 1) - FR(k, 1))) + tau * ny * (ul[k] - ur[k]);
    } else {
      fc[k + nf * 0] = 0.5 * (FL(k, 0) + FR(k, 0)) + pen * nx * (nx * (FL(k, 0) - FR(k, 0)) + ny * (FL(k, 1) - FR(k, 1)) + nz * (FL(k, 2) - FR(k, 2))) + tau * nx * (ul[k] - ur[k]);
      fc[k + nf * 1] = 0.5 * (FL(k, 1) + FR(k, 1)) + pen * ny * (nx * (FL(k, 0) - FR(k, 0)) + ny * (FL(k, 1) - FR(k, 1)) + nz * (FL(k, 2) - FR(k, 2))) + tau * ny * (ul[k] - ur[k]);
      fc[k + nf * 2] = 0.5 * (FL(k, 2) + FR(k, 2)) + pen * nz * (nx * (FL(k, 0) - FR(k, 0)) + ny * (FL(k, 1) - FR(k, 1)) + nz * (FL(k, 2) - FR(k, 2))) + tau * nz * (ul[k] - ur[k]);
    }
  }
#undef FL
#undef FR
  for (int k = 0; k < nf; k++) {
    fn[k] = fc[k + nf * 0] * n[0];
    for (int l = 1; l < nd; l++) fn[k] += fc[k + nf * l] * n[l];
  }
}

/* Ghost state of bdy_inters::set_inv_boundary_conditions (src/bdy_inters.cpp:391-
 * 420, 668-697, 884-889) for bctype 11, isothermal no-slip wall on a static mesh
 * (grid velocity zero): pressure extrapolated, T = T_wall, rho = p/(R_ref T). */
static void set_inv_bc(int nd, int type, const or_phys* ph, const double* ul, double* ur) {
  double gamma = ph->gamma, v_l[3], v_r[3] = {0., 0., 0.};
  double rho_l = ul[0];
  for (int i = 0; i < nd; i++) v_l[i] = ul[i + 1] / ul[0];
  double e_l = ul[nd + 1];
  double v_sq = 0.;
  for (int i = 0; i < nd; i++) v_sq += (v_l[i] * v_l[i]);
  double p_l = (gamma - 1.0) * (e_l - 0.5 * rho_l * v_sq);
  double p_r = p_l, rho_r = 0., e_r = 0.;
  if (type == 11) {
    double T_r = ph->T_wall;
    rho_r = p_r / (ph->R_ref * T_r);
    for (int i = 0; i < nd; i++) v_r[i] = 0.0;
    v_sq = 0.;
    for (int i = 0; i < nd; i++) v_sq += (v_r[i] * v_r[i]);
    e_r = (p_r / (gamma - 1.0)) + 0.5 * rho_r * v_sq;
  }
  ur[0] = rho_r;
  for (int i = 0; i < nd; i++) ur[i + 1] = rho_r * v_r[i];
  ur[nd + 1] = e_r;
}

/* ------------------------------------------------------------------------- */
/* CalcResidual steps */

#define UPT(a, j, i, k) (a)[(j) + (long)nu * ((i) + (long)N * (k))]
#define UPT4(a, j, i, k, l) (a)[(j) + (long)nu * ((i) + (long)N * ((k) + (long)nf * (l)))]

/* src/eles.cpp:1256-1327 */
static void extrapolate_solution(or_solver* s) {
  dgemm_ref(s->nfp, s->nf * s->n_eles, s->nu, 1.0, 0.0, s->opp_0, s->disu_upts[0], s->disu_fpts);
}

/* src/eles.cpp:1667-1766 */
static void calculate_gradient(or_solver* s) {
  long nuN = (long)s->nu * s->n_eles;
  for (int d = 0; d < s->nd; d++)
    dgemm_ref(s->nu, s->nf * s->n_eles, s->nu, 1.0, 0.0, s->opp_4[d], s->disu_upts[0], s->grad_disu_upts + nuN * s->nf * d);
}

/* src/eles.cpp:1331-1418 */
static void evaluate_invFlux(or_solver* s) {
  int nu = s->nu, N = s->n_eles, nf = s->nf, nd = s->nd;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double u[5], f[15];
    for (int j = 0; j < nu; j++) {
      for (int k = 0; k < nf; k++) u[k] = UPT(s->disu_upts[0], j, i, k);
      calc_invf(nd, s->ph.gamma, u, f);
      for (int k = 0; k < nf; k++)
        for (int l = 0; l < nd; l++) {
          double t = 0.;
          for (int m = 0; m < nd; m++) t += JG(l, m, j, i) * f[k + nf * m];
          UPT4(s->tdisf_upts, j, i, k, l) = t;
        }
    }
  }
}

/* src/int_inters.cpp:225-354 */
static void common_invFlux(or_solver* s) {
  int nf = s->nf, nd = s->nd, nfp = s->nfp, N = s->n_eles;
  long nfpN = (long)nfp * N, npts = (long)s->n_inters * s->nfpf;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts; q++) {
    int a = s->face_l[q], b = s->face_r[q];
    double ul[5], ur[5], fl[15], fr[15], n[3] = {0, 0, 0}, fn[5];
    for (int k = 0; k < nf; k++) { ul[k] = s->disu_fpts[a + nfpN * k]; ur[k] = s->disu_fpts[b + nfpN * k]; }
    for (int m = 0; m < nd; m++) n[m] = s->norm_fpts[a + nfpN * m];
    if (s->deck.riemann_solve_type == 2) {
      roe_flux(nd, s->ph.gamma, ul, ur, n, fn);
    } else {
      calc_invf(nd, s->ph.gamma, ul, fl);
      calc_invf(nd, s->ph.gamma, ur, fr);
      rusanov_flux(nd, s->ph.gamma, ul, ur, fl, fr, n, fn);
    }
    double tdl = s->tdA_fpts[a], tdr = s->tdA_fpts[b];
    for (int k = 0; k < nf; k++) {
      s->norm_tconf_fpts[a + nfpN * k] = fn[k] * tdl;
      s->norm_tconf_fpts[b + nfpN * k] = -fn[k] * tdr;
    }
    if (s->deck.viscous) {
      double pen = ldg_pen(nd, s->ph.pen_fact, n);
      for (int k = 0; k < nf; k++) {
        double uc = 0.5 * (ul[k] + ur[k]) - pen * (ul[k] - ur[k]);
        s->delta_disu_fpts[a + nfpN * k] = (uc - ul[k]);
        s->delta_disu_fpts[b + nfpN * k] = (uc - ur[k]);
      }
    }
  }
}

/* bdy_inters::evaluate_boundaryConditions_invFlux, src/bdy_inters.cpp:245-389:
 * central flux (convective_flux_boundary, src/inters.cpp:368-387) between the
 * trace and the ghost state, or the Roe flux with riemann_solve_type 2
 * (src/bdy_inters.cpp:330); LDG Dirichlet common solution u_c = (u_l + u_r)/2. */
static void bdy_invFlux(or_solver* s) {
  int nf = s->nf, nd = s->nd;
  long nfpN = (long)s->nfp * s->n_eles;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < s->n_bdy; q++) {
    int a = s->bdy_fpt[q];
    double ul[5], ur[5], fl[15], fr[15], n[3] = {0, 0, 0}, fn[5], fnl[5], fnr[5];
    for (int m = 0; m < nd; m++) n[m] = s->norm_fpts[a + nfpN * m];
    for (int k = 0; k < nf; k++) ul[k] = s->disu_fpts[a + nfpN * k];
    set_inv_bc(nd, s->bdy_type[q], &s->ph, ul, ur);
    calc_invf(nd, s->ph.gamma, ul, fl);
    calc_invf(nd, s->ph.gamma, ur, fr);
    for (int k = 0; k < nf; k++) {
      fnl[k] = 0.; fnr[k] = 0.;
      for (int l = 0; l < nd; l++) { fnl[k] += fl[k + nf * l] * n[l]; fnr[k] += fr[k + nf * l] * n[l]; }
    }
    for (int k = 0; k < nf; k++) fn[k] = 0.5 * (fnl[k] + fnr[k]);
    if (s->deck.riemann_solve_type == 2) roe_flux(nd, s->ph.gamma, ul, ur, n, fn);
    for (int k = 0; k < nf; k++) s->norm_tconf_fpts[a + nfpN * k] = fn[k] * s->tdA_fpts[a];
    if (s->deck.viscous)
      for (int k = 0; k < nf; k++) {
        double uc = 0.5 * (ur[k] + ul[k]);
        s->delta_disu_fpts[a + nfpN * k] = (uc - ul[k]);
      }
  }
}

/* src/eles.cpp:1770-1968 */
static void correct_gradient(or_solver* s) {
  int nu = s->nu, N = s->n_eles, nf = s->nf, nd = s->nd;
  long nuN = (long)nu * N;
  for (int d = 0; d < nd; d++)
    dgemm_ref(nu, nf * N, s->nfp, 1.0, 1.0, s->opp_5[d], s->delta_disu_fpts, s->grad_disu_upts + nuN * nf * d);
  double* g = s->grad_disu_upts;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < nu; j++) {
      double detjac = s->detjac_upts[j + nu * i];
      double inv_detjac = 1.0 / detjac;
      double rx = JG(0, 0, j, i) * inv_detjac, ry = JG(0, 1, j, i) * inv_detjac;
      double sx = JG(1, 0, j, i) * inv_detjac, sy = JG(1, 1, j, i) * inv_detjac;
      if (nd == 2) {
        for (int k = 0; k < nf; k++) {
          double ur = UPT4(g, j, i, k, 0), us = UPT4(g, j, i, k, 1);
          UPT4(g, j, i, k, 0) = ur * rx + us * sx;
          UPT4(g, j, i, k, 1) = ur * ry + us * sy;
        }
      } else {
        double rz = JG(0, 2, j, i) * inv_detjac, sz = JG(1, 2, j, i) * inv_detjac;
        double tx = JG(2, 0, j, i) * inv_detjac, ty = JG(2, 1, j, i) * inv_detjac, tz = JG(2, 2, j, i) * inv_detjac;
        for (int k = 0; k < nf; k++) {
          double ur = UPT4(g, j, i, k, 0), us = UPT4(g, j, i, k, 1), ut = UPT4(g, j, i, k, 2);
          UPT4(g, j, i, k, 0) = ur * rx + us * sx + ut * tx;
          UPT4(g, j, i, k, 1) = ur * ry + us * sy + ut * ty;
          UPT4(g, j, i, k, 2) = ur * rz + us * sz + ut * tz;
        }
      }
    }
}

/* src/eles.cpp:1973-2052 */
static void extrapolate_corrected_gradient(or_solver* s) {
  long nuN = (long)s->nu * s->n_eles, nfpN = (long)s->nfp * s->n_eles;
  for (int d = 0; d < s->nd; d++)
    dgemm_ref(s->nfp, s->nf * s->n_eles, s->nu, 1.0, 0.0, s->opp_6, s->grad_disu_upts + nuN * s->nf * d,
              s->grad_disu_fpts + nfpN * s->nf * d);
}

/* Eddy-viscosity SGS flux at a solution point, src/eles.cpp:2456-2778 (CPU path,
 * no wall model): sgs_model 0 Smagorinsky (C_s = 0.1), 1 WALE (C_w = 0.5,
 * eps = 1e-12); turbulent Prandtl number 0.5.  Kept as the reference CPU build
 * computes it, with its two index slips: the "velocity" is u(i) = temp_u(i)/rho for
 * i < d (src/eles.cpp:2483: density/rho in slot 0, the momenta shifted by
 * one), and du(i,j) = (d(rho u_j)/dx_i - u(j) d(rho)/dx_j)/rho takes the density
 * derivative along x_j (src/eles.cpp:2656).  u (nf), g (nf,nd) col-major, sgsf
 * (nf,nd) col-major physical flux. */
static void calc_sgsf(const or_solver* s, const double* u, const double* g, double detjac, double* sgsf) {
  int nd = s->nd, nf = s->nf;
  double diag = 0.0, Smod = 0.0, ke = 0.0, Pr = 0.5;
  double uu[3], drho[3], dene[3], dke[3], de[3], dmom[3][3], du[3][3], S[3][3];
  double rho = u[0];
  for (int i = 0; i < nd; i++) {
    uu[i] = u[i] / rho;
    ke += 0.5 * pow(uu[i], 2);
  }
  double inte = u[nf - 1] / rho - ke;
  for (int k = 0; k < nf * nd; k++) sgsf[k] = 0.0;
  /* calc_ele_vol: detjac * reference volume (quads 4, hexes 8, tets 8/6) */
  double vol = s->ele_type == 2 ? detjac * 8. / 6. : detjac * (nd == 3 ? 8. : 4.);
  double delta = s->deck.filter_ratio * pow(vol, 1. / nd) / (s->order + 1.);
  for (int i = 0; i < nd; i++) {
    drho[i] = g[0 + nf * i];
    dene[i] = g[nf - 1 + nf * i];
    for (int j = 1; j < nf - 1; j++) dmom[i][j - 1] = g[j + nf * i];
  }
  for (int i = 0; i < nd; i++) {
    dke[i] = ke * drho[i];
    for (int j = 0; j < nd; j++) {
      du[i][j] = (dmom[i][j] - uu[j] * drho[j]) / rho;
      dke[i] += rho * uu[j] * du[i][j];
    }
    de[i] = (dene[i] - dke[i] - drho[i] * inte) / rho;
  }
  for (int i = 0; i < nd; i++) {
    for (int j = 0; j < nd; j++) S[i][j] = (du[i][j] + du[j][i]) / 2.0;
    diag += S[i][i] / 3.0;
  }
  for (int i = 0; i < nd; i++) S[i][i] -= diag;
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++) Smod += 2.0 * S[i][j] * S[i][j];
  Smod = sqrt(Smod);
  double mu_t, Cs;
  if (s->deck.SGS_model == 0) {
    Cs = 0.1;
    mu_t = rho * Cs * Cs * delta * delta * Smod;
  } else {
    Cs = 0.5;
    double num = 0.0, denom = 0.0, eps = 1.e-12, Sq[3][3];
    diag = 0.0;
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) {
        Sq[i][j] = 0.0;
        for (int k = 0; k < nd; ++k) Sq[i][j] += (du[i][k] * du[k][j] + du[j][k] * du[k][i]) / 2.0;
        diag += du[i][j] * du[j][i] / 3.0;
      }
    for (int i = 0; i < nd; i++) Sq[i][i] -= diag;
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) {
        num += Sq[i][j] * Sq[i][j];
        denom += S[i][j] * S[i][j];
      }
    denom = pow(denom, 2.5) + pow(num, 1.25);
    num = pow(num, 1.5);
    mu_t = rho * Cs * Cs * delta * delta * num / (denom + eps);
  }
  for (int j = 0; j < nd; j++) {
    sgsf[0 + nf * j] = 0.0;
    sgsf[nf - 1 + nf * j] = -1.0 * s->ph.gamma * mu_t / Pr * de[j];
    for (int i = 1; i < nf - 1; i++) sgsf[i + nf * j] = -2.0 * mu_t * S[i - 1][j];
  }
}

/* SGS flux of a supplied kinematic eddy viscosity nu_t, in the form of the
 * reference's GPU build (SGS_flux_kernel, src/cuda_kernels.cu:1466-1660; velocity
 * and energy gradients of vis_NS_flux, :880-960): v_i = U_{i+1}/rho,
 * dv_i/dx_j = (dU_{i+1}/dx_j - v_i drho/dx_j)/rho, Sd the deviatoric strain,
 * de_j = (dE/dx_j - (E/rho) drho/dx_j)/rho - v_i dv_i/dx_j;
 * F(1+i, j) = -2 rho nu_t Sd(i,j), F(nf-1, j) = -gamma rho nu_t / 0.5 de_j. */
static void calc_sgsf_nu(const or_solver* s, const double* u, const double* g, double nu_t, double* sgsf) {
  int nd = s->nd, nf = s->nf;
  double rho = u[0], v[3], vg[3][3], de[3], Sd[3][3], tr = 0.0;
  for (int i = 0; i < nd; i++) v[i] = u[1 + i] / rho;
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++) vg[i][j] = (g[(1 + i) + nf * j] - v[i] * g[0 + nf * j]) / rho;
  for (int j = 0; j < nd; j++) {
    double t = (g[(nf - 1) + nf * j] - (u[nf - 1] / rho) * g[0 + nf * j]) / rho;
    for (int i = 0; i < nd; i++) t -= v[i] * vg[i][j];
    de[j] = t;
  }
  for (int i = 0; i < nd; i++) tr += vg[i][i];
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++) Sd[i][j] = 0.5 * (vg[i][j] + vg[j][i]) - (i == j ? tr / 3.0 : 0.0);
  double mu_t = rho * nu_t;
  for (int j = 0; j < nd; j++) {
    sgsf[0 + nf * j] = 0.0;
    for (int i = 0; i < nd; i++) sgsf[(1 + i) + nf * j] = -2.0 * mu_t * Sd[i][j];
    sgsf[(nf - 1) + nf * j] = -s->ph.gamma * mu_t / 0.5 * de[j];
  }
}

/* Replaces the metrics set_transforms computed (src/eles.cpp:4557-4847) with the
 * caller's, in the same layouts: detjac_upts (Nu,N), JGinv_upts (d,d,Nu,N),
 * tdA_fpts (Nfp,N), norm_fpts (Nfp,N,d).  The residual reads them as they are. */
void or_set_metrics(or_solver* s, const double* detjac, const double* JGinv, const double* tdA,
                    const double* norm) {
  const size_t nuN = (size_t)s->nu * s->n_eles, nfpN = (size_t)s->nfp * s->n_eles, nd = s->nd;
  memcpy(s->detjac_upts, detjac, sizeof(double) * nuN);
  memcpy(s->JGinv_upts, JGinv, sizeof(double) * nuN * nd * nd);
  memcpy(s->tdA_fpts, tdA, sizeof(double) * nfpN);
  memcpy(s->norm_fpts, norm, sizeof(double) * nfpN * nd);
}

void or_set_sgs_nu(or_solver* s, const double* nu_t) {
  if (!s->sgs_nu) return;
  memcpy(s->sgs_nu, nu_t, sizeof(double) * (size_t)s->nu * s->n_eles);
}

/* src/eles.cpp:2321-2453 (static mesh, no wall model) */
static void evaluate_viscFlux(or_solver* s) {
  int nu = s->nu, N = s->n_eles, nf = s->nf, nd = s->nd;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double u[5], g[15], f[15], sg[15];
    for (int j = 0; j < nu; j++) {
      for (int k = 0; k < nf; k++) {
        u[k] = UPT(s->disu_upts[0], j, i, k);
        for (int m = 0; m < nd; m++) g[k + nf * m] = UPT4(s->grad_disu_upts, j, i, k, m);
      }
      calc_visf(nd, &s->ph, u, g, f);
      if (s->deck.LES) {
        if (s->deck.SGS_model == 10) calc_sgsf_nu(s, u, g, s->sgs_nu[j + (long)nu * i], sg);
        else calc_sgsf(s, u, g, s->detjac_upts[j + (long)nu * i], sg);
        for (int k = 0; k < nf; k++)
          for (int l = 0; l < nd; l++) f[k + nf * l] += sg[k + nf * l];
        for (int k = 0; k < nf; k++)
          for (int l = 0; l < nd; l++) {
            UPT4(s->sgsf_upts, j, i, k, l) = 0.0;
            for (int m = 0; m < nd; m++) UPT4(s->sgsf_upts, j, i, k, l) += JG(l, m, j, i) * sg[k + nf * m];
          }
      }
      for (int k = 0; k < nf; k++)
        for (int l = 0; l < nd; l++)
          for (int m = 0; m < nd; m++) UPT4(s->tdisf_upts, j, i, k, l) += JG(l, m, j, i) * f[k + nf * m];
    }
  }
}

/* src/eles.cpp:3294-3368: SGS flux at the flux points, extrapolated (opp_0) */
static void evaluate_sgsFlux(or_solver* s) {
  long nuN = (long)s->nu * s->n_eles, nfpN = (long)s->nfp * s->n_eles;
  for (int d = 0; d < s->nd; d++)
    dgemm_ref(s->nfp, s->nf * s->n_eles, s->nu, 1.0, 0.0, s->opp_0, s->sgsf_upts + nuN * s->nf * d,
              s->sgsf_fpts + nfpN * s->nf * d);
}

/* src/eles.cpp:1423-1512 */
static void extrapolate_totalFlux(or_solver* s) {
  long nuN = (long)s->nu * s->n_eles;
  int cols = s->nf * s->n_eles;
  dgemm_ref(s->nfp, cols, s->nu, 1.0, 0.0, s->opp_1[0], s->tdisf_upts, s->norm_tdisf_fpts);
  for (int d = 1; d < s->nd; d++)
    dgemm_ref(s->nfp, cols, s->nu, 1.0, 1.0, s->opp_1[d], s->tdisf_upts + nuN * s->nf * d, s->norm_tdisf_fpts);
}

/* src/eles.cpp:1517-1589 */
static void calculate_divergence(or_solver* s) {
  long nuN = (long)s->nu * s->n_eles;
  int cols = s->nf * s->n_eles;
  dgemm_ref(s->nu, cols, s->nu, 1.0, 0.0, s->opp_2[0], s->tdisf_upts, s->div_tconf_upts);
  for (int d = 1; d < s->nd; d++)
    dgemm_ref(s->nu, cols, s->nu, 1.0, 1.0, s->opp_2[d], s->tdisf_upts + nuN * s->nf * d, s->div_tconf_upts);
}

/* src/int_inters.cpp:359-470 */
static void common_viscFlux(or_solver* s) {
  int nf = s->nf, nd = s->nd, nfp = s->nfp, N = s->n_eles;
  long nfpN = (long)nfp * N, npts = (long)s->n_inters * s->nfpf;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts; q++) {
    int a = s->face_l[q], b = s->face_r[q];
    double ul[5], ur[5], gl[15], gr[15], fl[15], fr[15], n[3] = {0, 0, 0}, fn[5];
    for (int k = 0; k < nf; k++) { ul[k] = s->disu_fpts[a + nfpN * k]; ur[k] = s->disu_fpts[b + nfpN * k]; }
    for (int m = 0; m < nd; m++)
      for (int l = 0; l < nf; l++) {
        gl[l + nf * m] = s->grad_disu_fpts[a + nfpN * (l + (long)nf * m)];
        gr[l + nf * m] = s->grad_disu_fpts[b + nfpN * (l + (long)nf * m)];
      }
    calc_visf(nd, &s->ph, ul, gl, fl);
    calc_visf(nd, &s->ph, ur, gr, fr);
    if (s->deck.LES) {
      /* the extrapolated (transformed) SGS flux is added to the physical viscous
         flux as the reference does (src/int_inters.cpp:415-427) */
      for (int m = 0; m < nd; m++)
        for (int l = 0; l < nf; l++) {
          fl[l + nf * m] += s->sgsf_fpts[a + nfpN * (l + (long)nf * m)];
          fr[l + nf * m] += s->sgsf_fpts[b + nfpN * (l + (long)nf * m)];
        }
    }
    for (int m = 0; m < nd; m++) n[m] = s->norm_fpts[a + nfpN * m];
    ldg_flux(nd, ul, ur, fl, fr, n, fn, s->ph.tau, s->ph.pen_fact);
    double tdl = s->tdA_fpts[a], tdr = s->tdA_fpts[b];
    for (int k = 0; k < nf; k++) {
      s->norm_tconf_fpts[a + nfpN * k] += fn[k] * tdl;
      s->norm_tconf_fpts[b + nfpN * k] += -fn[k] * tdr;
    }
  }
}

/* bdy_inters::evaluate_boundaryConditions_viscFlux, src/bdy_inters.cpp:903-1046,
 * flux_spec 1 (Dirichlet, bctype 11): f_c = f_l + tau n (u_l - u_r), with the SGS
 * traces added to f_l under LES; fn = f_c.n (src/inters.cpp:550-641). */
static void bdy_viscFlux(or_solver* s) {
  int nf = s->nf, nd = s->nd;
  long nfpN = (long)s->nfp * s->n_eles;
  double tau = s->ph.tau;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < s->n_bdy; q++) {
    int a = s->bdy_fpt[q];
    double ul[5], ur[5], gl[15], fl[15], fc[15] = {0}, n[3] = {0, 0, 0}, fn[5];
    for (int k = 0; k < nf; k++) ul[k] = s->disu_fpts[a + nfpN * k];
    for (int m = 0; m < nd; m++) n[m] = s->norm_fpts[a + nfpN * m];
    set_inv_bc(nd, s->bdy_type[q], &s->ph, ul, ur);
    for (int m = 0; m < nd; m++)
      for (int l = 0; l < nf; l++) gl[l + nf * m] = s->grad_disu_fpts[a + nfpN * (l + (long)nf * m)];
    calc_visf(nd, &s->ph, ul, gl, fl);
    if (s->deck.LES)
      for (int m = 0; m < nd; m++)
        for (int l = 0; l < nf; l++) fl[l + nf * m] += s->sgsf_fpts[a + nfpN * (l + (long)nf * m)];
    for (int k = 0; k < nf; k++)
      for (int m = 0; m < nd; m++) fc[k + nf * m] = fl[k + nf * m] + tau * n[m] * (ul[k] - ur[k]);
    for (int k = 0; k < nf; k++) {
      fn[k] = fc[k + nf * 0] * n[0];
      for (int l = 1; l < nd; l++) fn[k] += fc[k + nf * l] * n[l];
    }
    for (int k = 0; k < nf; k++) s->norm_tconf_fpts[a + nfpN * k] += fn[k] * s->tdA_fpts[a];
  }
}

/* src/eles.cpp:1594-1661 */
static void calculate_corrected_divergence(or_solver* s) {
  long n = (long)s->n_eles * s->nf * s->nfp;
  for (long i = 0; i < n; i++) s->norm_tconf_fpts[i] += -1.0 * s->norm_tdisf_fpts[i];
  dgemm_ref(s->nu, s->nf * s->n_eles, s->nfp, 1.0, 1.0, s->opp_3, s->norm_tconf_fpts, s->div_tconf_upts);
  long m = (long)s->n_eles * s->nf * s->nu;
  for (long i = 0; i < m; i++)
    if (isnan(s->div_tconf_upts[i])) { fprintf(stderr, "fr_oracle: NaN in residual\n"); break; }
}

/* src/solver.cpp:65-225 */
void or_residual(or_solver* s) {
  extrapolate_solution(s);
  if (s->deck.viscous) calculate_gradient(s);
  evaluate_invFlux(s);
  common_invFlux(s);
  bdy_invFlux(s);
  if (s->deck.viscous) {
    correct_gradient(s);
    extrapolate_corrected_gradient(s);
    evaluate_viscFlux(s);
  }
  if (s->deck.LES) evaluate_sgsFlux(s);
  extrapolate_totalFlux(s);
  calculate_divergence(s);
  if (s->deck.viscous) {
    common_viscFlux(s);
    bdy_viscFlux(s);
  }
  calculate_corrected_divergence(s);
}

/* eles::CalcIntegralQuantities, src/eles.cpp:6503-6643 (set_opp_volume_cubpts
 * :4111-4130 and the cubature-point detJ :5318-5352), for the state disu_upts(0) and
 * the corrected gradient grad_disu_upts of the last or_residual.  The cubature rule
 * (loc (d, ncub), weights) is an argument.  kinds: 0 kineticenergy, 1 vorticity,
 * 2 pressuredilatation, 3 straincolonproduct, 4 devstraincolonproduct. */
void or_integral_quantities(const or_solver* s, int ncub, const double* loc, const double* wts, int nq,
                            const int* kinds, double* out) {
  int nu = s->nu, N = s->n_eles, nf = s->nf, nd = s->nd;
  double* opp = (double*)xcalloc((size_t)ncub * nu, sizeof(double));
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < ncub; j++) opp[j + ncub * i] = nodal_basis(s, i, &loc[nd * j]);
  for (int m = 0; m < nq; m++) out[m] = 0.0;
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < ncub; j++) {
      double d[3][3], detjac, u[5], g[5][3], S[3][3];
      calc_d_pos(s, &loc[nd * j], i, d);
      if (nd == 2) detjac = d[0][0] * d[1][1] - d[0][1] * d[1][0];
      else
        detjac = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) - d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0])
               + d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
      for (int m = 0; m < nf; m++) {
        u[m] = 0.;
        for (int k = 0; k < nu; k++) u[m] += opp[j + ncub * k] * UPT(s->disu_upts[0], k, i, m);
      }
      for (int m = 0; m < nf; m++)
        for (int n = 0; n < nd; n++) {
          g[m][n] = 0.;
          for (int k = 0; k < nu; k++) g[m][n] += opp[j + ncub * k] * UPT4(s->grad_disu_upts, k, i, m, n);
        }
      double irho = 1. / u[0];
      double dudx = irho * (g[1][0] - u[1] * irho * g[0][0]);
      double dudy = irho * (g[1][1] - u[1] * irho * g[0][1]);
      double dvdx = irho * (g[2][0] - u[2] * irho * g[0][0]);
      double dvdy = irho * (g[2][1] - u[2] * irho * g[0][1]);
      double dudz = 0., dvdz = 0., dwdx = 0., dwdy = 0., dwdz = 0.;
      if (nd == 3) {
        dudz = irho * (g[1][2] - u[1] * irho * g[0][2]);
        dvdz = irho * (g[2][2] - u[2] * irho * g[0][2]);
        dwdx = irho * (g[3][0] - u[3] * irho * g[0][0]);
        dwdy = irho * (g[3][1] - u[3] * irho * g[0][1]);
        dwdz = irho * (g[3][2] - u[3] * irho * g[0][2]);
      }
      for (int m = 0; m < nq; m++) {
        double diagnostic = 0.0, tke = 0.0;
        if (kinds[m] == 0) {
          for (int n = 1; n < nf - 1; n++) tke += 0.5 * u[n] * u[n];
          diagnostic = irho * tke;
        } else if (kinds[m] == 1) {
          double wz = dvdx - dudy;
          diagnostic = wz * wz;
          if (nd == 3) {
            double wx = dwdy - dvdz, wy = dudz - dwdx;
            diagnostic += wx * wx + wy * wy;
          }
          diagnostic *= 0.5 / irho;
        } else if (kinds[m] == 2) {
          for (int n = 1; n < nf - 1; n++) tke += 0.5 * u[n] * u[n];
          double pressure = (s->ph.gamma - 1.0) * (u[nf - 1] - irho * tke);
          diagnostic = nd == 2 ? pressure * (dudx + dvdy) : pressure * (dudx + dvdy + dwdz);
        } else {
          S[0][0] = dudx;
          S[0][1] = (dudy + dvdx) / 2.0;
          S[1][0] = S[0][1];
          S[1][1] = dvdy;
          double diag = (S[0][0] + S[1][1]) / 3.0;
          if (nd == 3) {
            S[0][2] = (dudz + dwdx) / 2.0;
            S[1][2] = (dvdz + dwdy) / 2.0;
            S[2][0] = S[0][2];
            S[2][1] = S[1][2];
            S[2][2] = dwdz;
            diag += S[2][2] / 3.0;
          }
          if (kinds[m] == 4)
            for (int a = 0; a < nd; a++) S[a][a] -= diag;
          for (int a = 0; a < nd; a++)
            for (int b = 0; b < nd; b++) diagnostic += S[a][b] * S[a][b];
        }
        out[m] += diagnostic * wts[j] * detjac;
      }
    }
  }
  free(opp);
}

/* src/eles.cpp:1003-1198, dt_type 0 */
void or_advance(or_solver* s, int stage, int adv_type) {
  int nu = s->nu, N = s->n_eles, nf = s->nf;
  double dt = s->ph.dt, cs = s->ph.const_src;
  if (adv_type == 0) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf; i++)
      for (int ic = 0; ic < N; ic++)
        for (int inp = 0; inp < nu; inp++)
          UPT(s->disu_upts[0], inp, ic, i) -= dt * (UPT(s->div_tconf_upts, inp, ic, i) / s->detjac_upts[inp + nu * ic] - cs - UPT(s->src_upts, inp, ic, i));
    return;
  }
  static const double rk4a[5] = {0.0, -0.417890474499852, -1.192151694642677, -1.697784692471528, -1.514183444257156};
  static const double rk4b[5] = {0.149659021999229, 0.379210312999627, 0.822955029386982, 0.699450455949122, 0.153057247968152};
  double a = rk4a[stage], b = rk4b[stage];
#pragma omp parallel for schedule(static)
  for (int ic = 0; ic < N; ic++)
    for (int i = 0; i < nf; i++)
      for (int inp = 0; inp < nu; inp++) {
        double rhs = -UPT(s->div_tconf_upts, inp, ic, i) / s->detjac_upts[inp + nu * ic] + cs + UPT(s->src_upts, inp, ic, i);
        double res = UPT(s->disu_upts[1], inp, ic, i);
        res = a * res + dt * rhs;
        UPT(s->disu_upts[1], inp, ic, i) = res;
        UPT(s->disu_upts[0], inp, ic, i) += b * res;
      }
}

/* set_ics, ic_form 7 (Taylor-Green), src/eles.cpp:434-470, 555-568 */
void or_set_ics_tgv(or_solver* s, double rho) {
  int nu = s->nu, N = s->n_eles, nf = s->nf;
  double gamma = s->ph.gamma;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < nu; j++) {
      double pos[3];
      for (int k = 0; k < 3; k++) pos[k] = s->pos_upts[j + nu * (i + (long)N * k)];
      double ics[5];
      double p = 100 + rho / 16.0 * (cos(2.0 * pos[0]) + cos(2.0 * pos[1])) * (cos(2.0 * pos[2]) + 2.0);
      ics[0] = rho;
      ics[1] = sin(pos[0]) * cos(pos[1]) * cos(pos[2]);
      ics[2] = -1.0 * cos(pos[0]) * sin(pos[1]) * cos(pos[2]);
      ics[3] = 0.0;
      ics[4] = p / (gamma - 1.0) + 0.5 * rho * (ics[1] * ics[1] + ics[2] * ics[2] + ics[3] * ics[3]);
      for (int k = 0; k < nf; k++) UPT(s->disu_upts[0], j, i, k) = ics[k];
    }
}

/* src/HiFiLES.cpp:154-187 */
void or_run_steps(or_solver* s, int n_steps) {
  int nst = s->deck.adv_type == 3 ? 5 : 1;
  for (int it = 0; it < n_steps; it++)
    for (int st = 0; st < nst; st++) {
      or_residual(s);
      or_advance(s, st, s->deck.adv_type);
    }
}

/* src/output.cpp:1452-1522 + src/eles.cpp:6086-6116 */
void or_norm_residual(const or_solver* s, int norm_type, double* out) {
  int nu = s->nu, N = s->n_eles, nf = s->nf;
  for (int k = 0; k < 5; k++) out[k] = 0.0;
  for (int f = 0; f < nf && f < 5; f++) {
    double sum = 0.;
    for (int i = 0; i < N; i++) {
      double cell = 0.;
      for (int j = 0; j < nu; j++) {
        double r = UPT(s->div_tconf_upts, j, i, f) / s->detjac_upts[j + nu * i] - s->ph.const_src - UPT(s->src_upts, j, i, f);
        if (norm_type == 0) cell = fmax(cell, fabs(r));
        else if (norm_type == 1) cell += fabs(r);
        else cell += r * r;
      }
      if (norm_type == 0) sum = fmax(cell, sum);
      else sum += cell;
    }
    long n_upts = (long)N * nu;
    if (norm_type == 0) out[f] = sum;
    else if (norm_type == 1) out[f] = sum / n_upts;
    else out[f] = sqrt(sum) / n_upts;
  }
}
