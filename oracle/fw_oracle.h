/*
 * fw_oracle.h - CPU restatement of the ADVENTURE_FullWave HDDM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  The product path never links or calls it.
 *
 * Every function restates one reference routine (file:line given at each
 * declaration) in plain C99, following the reference's arithmetic
 * (struct-complex multiply / naive reciprocal of lib/std/complex_adv.c,
 * C99 double _Complex in the skyline code) and loop order.  The reference
 * itself cannot be compiled here (every source includes the un-vendored
 * ADVENTURE_IO header Adv/AdvDocument.h), so parity of this restatement is
 * pinned by closed forms, invariants and the reference's shipped
 * convergence log instead of by linking the reference.
 */
#ifndef FW_ORACLE_H
#define FW_ORACLE_H

#include "../include/fwh_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference Complex (lib/std/types.h:15-18): {re, im} */
typedef struct { double re, im; } fwo_c;

/* ---- element kernel (a2) ---------------------------------------------- */
/* tet_MakeElement_EM_E (lib/fem/make_elem_em.c:14-400) for one element:
 * ae[6*6] (symmetric, ae[i*6+j] couples local edges i and j) and be[6].
 * joe: 4 vertices x 3 complex components (re,im) = 24 doubles, used when
 * src != 0.  pml: NULL or 6 doubles (complex sx, sy, sz).               */
int fwo_element(const double *xyz12, double eps, double sigma, double omega,
                const double *pml, int src, const double *joe24,
                double *ae72, double *be12);

/* Delta-gap load (make_elem_em.c:320-396): be[6] of one element with the
 * 4 vertex gap voltages dv4 (0 outside every gap box).                    */
int fwo_dg_load(const double *xyz12, double sigma, double omega, const double *dv4,
                double *be12);

/* ---- element post-processing ------------------------------------------- */
/* HDDM_post_em_E (main/fullwave/fullwave.c:640-763) for one element:
 * xe12 = solution at the 6 edges (re,im), nu = 1/mu of the element's
 * material (0 when its flag is not a listed material).  rot6, e6, j6 get
 * 3 complex each: i (nu/omega) rot E, 2 (nu/omega) E(g), sigma (nu/omega)
 * E(g) in the reference's scalings.  Returns <0 for a degenerate element. */
int fwo_post_element(const double *xyz12, const double *xe12, double nu,
                     double sigma, double omega, double *rot6, double *e6,
                     double *j6);

/* ---- skyline LDL^T (a4-a6) on a dense-described profile --------------- */
/* advlas_skyline_ldl_decomposite (lib/hddm/advlas_skyline.c:79-128):
 * mat (2*nlen doubles) holds rows i = 0..n-1, row i columns idx1[i]..i at
 * position idx2[i]+j.  Factor in place (L strict lower, 1/d on diagonal). */
void fwo_sky_factor(int n, const int *idx1, const int *idx2, double *mat);
/* advlas_skyline_ldl_solve (lib/hddm/advlas_skyline.c:134-167), in place */
void fwo_sky_solve(int n, const int *idx1, const int *idx2, const double *mat,
                   double *v);

/* ---- per-subdomain pipeline (a2-a7) ----------------------------------- */
/* Opaque per-subdomain state built from a mesh description.              */
typedef struct fwo_model fwo_model;

/* Assemble every subdomain (HDDM_make_TMP_AIJ_em_E, main/fullwave/
 * fullwave.c:221-259), apply the Dirichlet structure and build + factor
 * the skyline (hddm_solver.h:485-514).  Returns NULL on failure.         */
fwo_model *fwo_model_build(const fwh_mesh_desc *m);
void fwo_model_free(fwo_model *md);

/* Sizes of subdomain s: n (edge DOFs), nlen (skyline entries), nnz of the
 * assembled lower triangle, number of Dirichlet entries.                 */
void fwo_model_sizes(const fwo_model *md, int s, int *n, int *nlen, int *nnz,
                     int *nbc);
/* Copy out subdomain s: assembled lower-triangle CSR (rowptr n+1, col nnz,
 * val 2*nnz), rhs f (2n), skyline idx1/idx2 (n each) and factor (2*nlen). */
void fwo_model_get(const fwo_model *md, int s, int *rowptr, int *col,
                   double *val, double *f, int *idx1, int *idx2,
                   double *factor);

/* Schur step of one subdomain (hddm_solver.h:559-622 with
 * hddm_solver_sp.h:58-69): interface values vbc (2*ninbc doubles, in the
 * subdomain's interface order), returns y = A x (2n doubles) where
 * x solves the Dirichlet problem with zero load, and x itself.           */
void fwo_model_schur_sub(fwo_model *md, int s, const double *vbc, double *x,
                         double *y);

/* ---- full s-mode HDDM solve (a7, a8, a11) ----------------------------- */
typedef struct {
  double hddm_eps;      /* max-norm ratio stop (-hddm-eps)                 */
  int    hddm_max_loop; /* (-hddm-max-loop)                                */
  int    dom_solver;    /* 0 = skyline LDL^T (shipped), 1 = COCG + ICC     */
  double sol_eps;       /* subdomain COCG stop (-l-sol-eps)                */
  double ic_shift;      /* ICC acceleration (-l-sol-ic-shift)              */
  int    max_sub;       /* stop after this many subdomain solves (<=0: no) */
  int    perturb;       /* 0: reference summation order; 1: interface dot
                           products accumulated in reverse unknown order
                           (measures rounding sensitivity only)            */
  int    nthreads;      /* > 1: the main loop's subdomain Schur steps on
                           that many OpenMP threads, accumulated afterwards
                           in the reference's order (the same sums)        */
} fwo_opts;

typedef struct {
  int    loops;         /* hddm_sw.loop at exit                            */
  int    iters;         /* main-loop interface iterations executed         */
  int    converged;     /* 1 if the stop test fired                        */
  double r0;            /* ||r0||_inf                                      */
  double seconds_loop;  /* wall time of the main loop                      */
  long   sub_solves;    /* subdomain solves performed                      */
} fwo_stats;

/* hddm_solver.h:59-183 in s-mode.  resid: [max_loop+1] residual history
 * (entry 0 = 1.0).  x_out: per subdomain solution, concatenated, 2*n_s
 * doubles each.  u_out: per part owned interface values, concatenated,
 * 2*infree each.  Either output may be NULL.  Returns 0 on success.      */
int fwo_hddm_solve(const fwh_mesh_desc *m, const fwo_opts *o, double *resid,
                   double *x_out, double *u_out, fwo_stats *st);

/* fwo_hddm_solve with the load replaced: load holds every subdomain's
 * per-DOF load (2*n_s doubles each, concatenated in subdomain order, the
 * layout of x_out); an adjoint solve A lambda = g uses it (A^T = A).      */
int fwo_hddm_solve_load(const fwh_mesh_desc *m, const fwo_opts *o, const double *load,
                        double *resid, double *x_out, double *u_out, fwo_stats *st);

/* Time `n_iter` main-loop interface iterations on a built model (the
 * CPU baseline) with `nthreads` OpenMP threads over the subdomain Schur
 * steps (accumulated in the reference's order); returns seconds.         */
double fwo_hddm_time_iterations(const fwh_mesh_desc *m, int n_iter,
                                long *sub_solves, int nthreads);

/* ---- subdomain COCG + ICC (a9) ----------------------------------------- */
/* mymt_CS_AIJ_Solver_COCG (aij_solver_cg.h:9-44, 113-355) with ICC
 * (aij_pc_func.h:36-167) on a lower-triangle CSR system (row's diagonal
 * last).  x (2n) holds the initial guess and receives the solution.
 * Returns the reference status (>=0 iterations, <0 error code).          */
int fwo_cocg_icc(int n, const int *rowptr, const int *col, const double *val,
                 const double *b, double *x, double eps, double ic_shift,
                 int use_icc, double *resid_hist);

/* IC(0) factor only (aij_pc_func.h:36-108): out has 2*(nnz+n) doubles.    */
void fwo_icc_make(int n, const int *rowptr, const int *col, const double *val,
                  double ic_shift, double *pc);

#ifdef __cplusplus
}
#endif

#endif /* FW_ORACLE_H */
