/*
 * fwhip.h - C ABI of the MI355X (gfx950) HDDM full-wave library libfwhip.so.
 *
 * Plain C linkage, plain pointers and sizes; no HIP or torch types cross this
 * boundary.  Complex values are interleaved doubles (re, im), the layout of
 * the reference's Complex struct (lib/std/types.h:15-18) and of C99
 * double _Complex.  Every entry point returns 0 on success or a negative
 * FWH_E_* code; fwh_last_error() returns the message of the last failure in
 * the calling thread.  The reference turns failures into
 * emessage(code, abort, str) -> COM_Exit (lib/std/errormessage.h:13); a
 * reference-side hook does the same with these codes.
 *
 * Each entry point names the reference interface it replaces.  The reference
 * plugs its subdomain work into the HDDM_FUNC hook table
 * (lib/std/types.h:629-730) filled by set_hddm_func_fullwave
 * (main/fullwave/fullwave.c:23-146) and driven by hddm_solver_CS
 * (lib/hddm/hddm_solver.h:25-183); this library keeps that call structure
 * (first step / iterations / last step) but runs every subdomain of every
 * part on one device.
 */
#ifndef FWHIP_H
#define FWHIP_H

#include "fwh_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FWH_OK 0
#define FWH_E_ARG (-1)     /* invalid argument / inconsistent model        */
#define FWH_E_HIP (-2)     /* HIP runtime failure                           */
#define FWH_E_MESH (-3)    /* degenerate element (six-times volume <= 0,
                              zero-length edge: emessage 2050 in
                              lib/fem/tetrahedral.c:257-258)                */
#define FWH_E_STATE (-4)   /* call out of order (e.g. iterate before load)  */
#define FWH_E_LIMIT (-5)   /* a size exceeds a compiled kernel limit        */

typedef struct fwh_ctx fwh_ctx;

/* Solver options: the reference's HDDM_Opt / SolvOpt fields on this path
 * (lib/std/types.h:116-141; defaults lib/std/defs.h:120-205).            */
typedef struct fwh_opts {
  double hddm_eps;      /* -hddm-eps: stop when ||r||inf/||r0||inf < eps   */
  int    hddm_max_loop; /* -hddm-max-loop (0: not executed)                */
  int    plain;         /* 1: one subdomain per launch, reference order;
                           0: batched (all subdomains per launch, graph)   */
  int    graph_iters;   /* batched: iterations per captured graph (0=no
                           graph, plain launches)                          */
  int    schur;         /* 0: implicit Schur step, Dirichlet solves through
                           the factors every iteration (the reference's
                           form); 1: explicit S_s formed once from the
                           factors (fwh_schur_form), one streamed matvec
                           per iteration                                   */
} fwh_opts;

typedef struct fwh_stats {
  int    loops;         /* hddm_sw.loop at exit                            */
  int    iters;         /* interface iterations executed                   */
  int    converged;     /* stop test fired                                 */
  double r0;            /* ||r0||inf                                       */
  double final_resid;   /* last ||r||inf/||r0||inf                         */
  double ms_first;      /* first step (solve + PC) wall ms                 */
  double ms_loop;       /* main loop wall ms                               */
  double ms_last;       /* last step wall ms                               */
  long   sub_solves;    /* subdomain LDL^T solves performed                */
} fwh_stats;

typedef struct fwh_info {
  int    nsub, npart;
  int    n_iface;       /* global interface unknowns N_Gamma               */
  long   n_dof;         /* sum of subdomain DOFs                           */
  long   n_slots;       /* sum of subdomain interface DOFs                 */
  long   nnz;           /* sum of assembled lower-triangle entries         */
  long   nlen;          /* sum of skyline lengths                          */
  long   n_elem;        /* elements                                        */
  long   n_cpl;         /* interior-interface coupling entries (A_I,G)     */
  long   n_grow;        /* entries of the interface rows (A_G,.)           */
  int    max_n;         /* largest subdomain                               */
  double ms_assemble;   /* setup timings                                   */
  double ms_factor;
  double bytes_solve;   /* algorithmic bytes of one LDL^T solve of every
                           subdomain: sum 16 nlen + 40 n                   */
  double bytes_iter;    /* algorithmic bytes of one interface iteration    */
  double flops_factor;  /* complex-FP64 flops of the factorisation         */
  double ms_schur_form; /* explicit Schur complements: formation wall ms   */
  double smat_bytes;    /* their storage (16x16 tiles of the lower half)   */
  double bytes_smv;     /* algorithmic bytes of one explicit Schur step:
                           sum 8 ng(ng+1) + 36 ng                          */
  int    solve_buckets; /* size buckets of the batched LDL^T kernel        */
  int    solve_bucketed;/* 1: launched per bucket (faster at load), 0: one
                           launch                                          */
  double ms_solve_single, ms_solve_bucketed; /* the timings of that choice */
  int    smv_waves;     /* waves per workgroup of the explicit Schur matvec,
                           picked by timing when the S_s are formed        */
  double ms_smv_tune[3];/* its timings for 2, 4, 8 waves                   */
  int    solve_depth;   /* loads per lane in flight (8 or 16) of the
                           LDL^T kernel in the iteration sweeps, picked by
                           timing at load                                  */
  int    unit_depth;    /* the same for the Schur-complement formation     */
  double ms_solve_depth[2]; /* best batched solve time at depth 8 (0 when
                           a profile row is wider than 512), 16            */
  int    nsys;          /* systems of the batch (fwh_load_batch); the byte
                           and flop counts above cover all of them          */
  int    smv_order;     /* work order of the explicit Schur matvec picked
                           by timing: 0 largest skyline first, 1 most tiles
                           first, 2 most tiles first snaking over the CUs  */
  double ms_smv_order[3]; /* its timings                                   */
  int    schur_gemm;    /* 1: S_s formed as A_GG - W^T D^-1 W on the matrix
                           cores (W by blocked forward substitution); 0: by
                           the Schur step on every interface unit vector   */
  int    factor_blocked; /* 1: LDL^T factors by 16-row blocks on the matrix
                           cores; 0: one row at a time (FWH_FACTOR=row)     */
  double ms_host_index; /* host work of the load: patterns, assembly and
                           gather lists, couplings, profiles               */
  double ms_load;       /* the whole load (fwh_load / fwh_load_batch):
                           host work, uploads, assembly, factorisation and
                           the load-time kernel tuning                     */
  double ms_factor_kernel; /* device time (HIP events) of the LDL^T factor
                           kernel alone, all systems                       */
  double ms_schur_fwd;  /* device time of the forward substitution
                           W = L^-1 (-A_IG) of the explicit formation      */
  double ms_schur_gemm; /* device time of its tile products
                           S_s -= W^T D^-1 W                               */
  double flops_schur_fwd;  /* complex-FP64 flops of that substitution:
                           sum 8 ng (nlen - n)                             */
  double flops_schur_gemm; /* and of the tile products: sum over subdomains
                           8 nI ng (ng+1)/2 + 6 nI ng (nI interior rows)   */
  int    smv_split;     /* explicit matvec: 0 one work item per subdomain;
                           m > 0 subdomains split into items of at most
                           total tiles / (nCU m) tiles (picked by timing)  */
  int    smv_nwork;     /* its work items                                  */
  double ms_smv_split[4]; /* its timings for m = 0, 2, 4, 8 (-1: no
                           subdomain large enough to split)                */
  int    solve_waves;   /* waves per subdomain (1, 2 or 4) of the LDL^T
                           kernel outside the formation, picked by timing
                           at load                                         */
  double ms_solve_waves[3]; /* best batched solve time with 1, 2 and 4
                           waves at the chosen depth (-1: does not fit)    */
  int    smv_persist;   /* 1: the explicit matvec runs as a persistent
                           grid of smv_pgrid workgroups that claim the work
                           items from a queue (same items, same results);
                           0: one workgroup per item.  Picked by timing    */
  int    smv_pgrid;     /* workgroups of the persistent grid (resident
                           capacity: occupancy x CUs)                      */
  double ms_smv_persist[2]; /* its timings: one workgroup per item,
                           persistent (-1: not timed)                      */
  int    solve_persist; /* the batched LDL^T kernel outside the formation:
                           0 one workgroup per subdomain (solve_depth,
                           solve_waves); 1 persistent, workgroups of one
                           wave claiming subdomains from a queue; 2
                           persistent with two size-bucketed queues (the
                           solve_nlarge largest subdomains by a workgroup
                           of four waves, the others one wave each).
                           Picked by timing at load                        */
  int    solve_pgrid;   /* workgroups of the persistent grid               */
  int    solve_nlarge;  /* subdomains of the large queue (form 2)          */
  double ms_solve_persist[2]; /* best LDL^T solve time of forms 1 and 2
                           (-1: does not fit)                              */
  int    factors_resident; /* 1: the skyline factors are in device memory;
                           0: released (fwh_release_factors), recomputed
                           chunk by chunk by the solves that need them     */
  int    factor_chunks; /* chunks of that recomputation                    */
  double factor_bytes_released; /* device bytes the release freed          */
  double factor_chunk_bytes;    /* bytes of the recomputation buffers      */
  double factor_need_max;  /* recomputation bytes of the largest subdomain
                           (factor entries, block inverses, the factor
                           kernel's temporaries): the smallest budget      */
  double factor_need_total; /* the same summed over all subdomains         */
  double ms_solve_wide; /* best batched solve time of the four-wave form
                           over the wide row blocks (<= 2,048 entries,
                           depth 32; -1: the blocks are not built)         */
} fwh_info;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the reference has no device).                  */
int  fwh_create(fwh_ctx **ctx, int device);
void fwh_destroy(fwh_ctx *ctx);
const char *fwh_last_error(void);

/* ---- model load: assembly + Dirichlet structure + skyline + factor ------
 * Replaces, for every subdomain of every part, HDDM_make_TMP_AIJ_em_E
 * (main/fullwave/fullwave.c:221-259 -> tet_MakeElement_EM_E,
 * lib/fem/make_elem_em.c:14-400, mymt_CS_AIJ_Set/sort/DeleteZero) and the
 * factorisation half of hddm_solver_CS_solve_dom_1st_d
 * (lib/hddm/hddm_solver.h:485-514: Copy_TMP_AIJ, SetBC,
 * advlas_copy_nzindex_to_skyindex, advlas_copy_nzmat_to_skymat,
 * advlas_skyline_ldl_decomposite).  The model description is copied. */
int  fwh_load(fwh_ctx *ctx, const fwh_mesh_desc *m);
int  fwh_get_info(const fwh_ctx *ctx, fwh_info *info);

/* ---- batches of systems on one mesh (no reference counterpart: the
 * reference runs one hddm_solver_CS per model, main/fullwave/main.c:105) --
 * fwh_load_batch loads nsys systems that share the mesh and the
 * decomposition of descs[0] (every index array of descs[k] equal to
 * descs[0]'s) and have their own omega, eps, sigma, source and PML: a
 * frequency sweep is the typical batch.  Each system gets its own factors
 * (and Schur complements), its own interface COCG, stop test and status;
 * every batched launch of fwh_solve serves all systems (one grid row per
 * system), and the solve ends once every system's status is LAST.  Results
 * match nsys separate single-system solves bit for bit.  fwh_load is
 * fwh_load_batch with nsys = 1.
 * fwh_select chooses the system that the per-system calls address:
 * fwh_get_stats, fwh_get_x, fwh_get_u, fwh_get_resid, fwh_post,
 * fwh_set_load, fwh_grad_eps, fwh_schur_apply, fwh_ldl_solve_all and
 * fwh_get_sub (0 after a load).                                           */
int  fwh_load_batch(fwh_ctx *ctx, const fwh_mesh_desc *descs, int nsys);
int  fwh_select(fwh_ctx *ctx, int sys);

/* ---- the HDDM solve -----------------------------------------------------
 * fwh_solve replaces __HDDM_SOLVER_PARENT__ in s-mode
 * (lib/hddm/hddm_solver.h:59-183 with hddm_solver_s.h:7-438,
 * hddm_solver_sp.h:8-69, hddm_pc.h:52-137): LOOP_1ST, the main loop with
 * its status machine, LOOP_LAST.  The three steps are also exported:
 *   fwh_first   = LOOP_1ST + iteration_s_1st (hddm_solver_s.h:7-52,158-256)
 *   fwh_iterate = n x (LOOP + iteration_s)   (hddm_solver_s.h:69-112,260-400)
 *   fwh_last    = LOOP_LAST + get_result_sp  (hddm_solver_s.h:116-154,
 *                 main/fullwave/fullwave.c:500-514)                       */
int  fwh_solve(fwh_ctx *ctx, const fwh_opts *opts, fwh_stats *stats);
int  fwh_first(fwh_ctx *ctx, const fwh_opts *opts);
/* runs up to n iterations; *done = 1 once the status is LAST             */
int  fwh_iterate(fwh_ctx *ctx, const fwh_opts *opts, int n, int *done);
int  fwh_last(fwh_ctx *ctx);
int  fwh_get_stats(const fwh_ctx *ctx, fwh_stats *stats);

/* Results.  x: per subdomain solution (dm->x_c), concatenated in subdomain
 * order, 2*nmp doubles each.  u: interface values of the owned DOFs, part
 * by part in ln_part order (HDDM_VEC.u, 2*infree each).  resid: residual
 * history ||r||inf/||r0||inf for loops 0..nresid-1 (the values the
 * reference writes to calc_log/log_g_HDDM_Full_Wave).                    */
int  fwh_get_x(const fwh_ctx *ctx, double *x);
int  fwh_get_u(const fwh_ctx *ctx, double *u);
int  fwh_get_resid(const fwh_ctx *ctx, double *resid, int nresid);

/* Element post-processing of the last solution: HDDM_post_dom_em_E
 * (main/fullwave/fullwave.c:607-635) with HDDM_post_em_E (:640-763) for
 * every element, in model element order.  nu: per element 1/mu of its
 * material (7.957747e5 when the element's flag is a listed material, 0
 * otherwise, :670-675), or NULL for 7.957747e5 everywhere.  B, E, J receive
 * 3 complex values (6 doubles) per element: the reference's el_B_c
 * = i (nu/omega) rot E, el_E_c = 2 (nu/omega) E(centroid) and
 * el_Je_c = sigma (nu/omega) E(centroid), with the reference's scalings.  */
int  fwh_post(fwh_ctx *ctx, const double *nu, double *B, double *E, double *J);

/* ---- adjoint sensitivities (no reference counterpart) --------------------
 * fwh_set_load: the per-DOF load of the next solve (2*n_dof doubles in
 * fwh_get_x's layout; the global load is the sum over a DOF's copies), NULL
 * restores the assembled load.  Factors, Schur complements and the
 * preconditioner are reused, so an adjoint solve (A = A^T) costs one more
 * interface solve.
 * fwh_grad_eps: dJ/deps_e = 2 omega Re(lambda_e^T M_e E_e) for every element
 * (M_e: the element's edge mass matrix, PML-weighted when stretched; eps
 * absolute), from a forward solution E and the adjoint solution lambda of
 * A lambda = dJ/dE, both in fwh_get_x's layout; grad: n_elem doubles.      */
int  fwh_set_load(fwh_ctx *ctx, const double *f);
int  fwh_grad_eps(fwh_ctx *ctx, const double *xE, const double *xL, double *grad);

/* ---- single operators (parity tests, benchmarks) ----------------------- */
/* One Schur-complement application q = S p over all subdomains
 * (renewal_bc_sp + solve_dom_d + COM_res_sp + gather,
 * hddm_solver_sp.h:28-69, hddm_solver.h:559-622, hddm_solver_s.h:404-424).
 * p, q: owned interface values, 2*n_iface doubles, part-major.
 * mode: 0 implicit batched, 1 implicit plain (one subdomain per launch),
 * 2 explicit batched, 3 explicit plain (2, 3 form S_s first if needed).  */
int  fwh_schur_apply(fwh_ctx *ctx, const double *p, double *q, int mode);
/* Form the explicit subdomain Schur complements S_s = A_GG - A_GI A_II^-1
 * A_IG from the factors: the Schur step of solve_dom_d
 * (lib/hddm/hddm_solver.h:559-622) applied to every interface unit vector.
 * Idempotent; fwh_first forms them when opts.schur = 1.                   */
int  fwh_schur_form(fwh_ctx *ctx);
/* Release the skyline factors and block inverses after the Schur
 * complements are formed (the device memory of the factors, 16 bytes per
 * skyline entry and system).  No reference counterpart: the reference keeps
 * dm->dmat for the whole solve (hddm_solver.h:512-523).  The explicit
 * iteration needs no factors; the solves that do (fwh_first, fwh_last,
 * fwh_ldl_solve_all, the adjoint's solves through fwh_set_load) recompute
 * them chunk by chunk, consecutive subdomains whose factors fit
 * budget_bytes, with the load's kernels (same factors and pivots).  After
 * it the implicit Schur step is refused (FWH_E_STATE); a new load restores
 * resident factors.                                                       */
int  fwh_release_factors(fwh_ctx *ctx, long long budget_bytes);
/* Device bytes the skyline factors may take (all systems, with the block
 * inverses of the solve kernel; 0, the default: no limit).  Read by the
 * next fwh_load / fwh_load_batch: a model whose factors exceed it is loaded
 * streamed -- the factors are never all resident: the Schur complements are
 * formed chunk by chunk right after each chunk's factors are computed (in
 * buffers of at most budget_bytes), and the solves that need the factors
 * recompute them as after fwh_release_factors (the LDL^T kernel runs at
 * depth 16 with one wave per subdomain; the implicit Schur step is refused).
 * No reference counterpart: the reference keeps every subdomain's factor in
 * host memory (hddm_solver.h:512).                                         */
int  fwh_set_factor_budget(fwh_ctx *ctx, long long budget_bytes);
/* advlas_skyline_ldl_solve (lib/hddm/advlas_skyline.c:134-167) applied to
 * every subdomain at once: rhs/x are 2*n_dof doubles (in place).          */
int  fwh_ldl_solve_all(fwh_ctx *ctx, double *rhs, int plain);
/* One skyline matrix in the reference's DMatrix layout (lib/std/types.h:
 * 143-157): n rows, row i holding columns idx1[i]..i at mat[idx2[i] + j]
 * (2 doubles each), rows contiguous.  fwh_sky_factor replaces
 * advlas_skyline_ldl_decomposite (lib/hddm/advlas_skyline.c:79-128): mat is
 * overwritten by L (strictly lower) with 1/d on the diagonal; dinv (2n
 * doubles, may be NULL) receives 1/d.  fwh_sky_solve replaces
 * advlas_skyline_ldl_solve (:134-167): v (2n doubles) in place, mat the
 * factor.  Both stage the matrix on the device and run the batched kernels
 * on it (pivot-free complex-symmetric LDL^T, Smith reciprocals).         */
int  fwh_sky_factor(fwh_ctx *ctx, int n, const int *idx1, const int *idx2,
                    double *mat, double *dinv);
int  fwh_sky_solve(fwh_ctx *ctx, int n, const int *idx1, const int *idx2,
                   const double *mat, double *v);
/* Device time (HIP events on the library stream) of one batched LDL^T solve
 * of all subdomains, of one batched implicit Schur-step kernel (Dirichlet
 * load, solve, interface rows of A x) and of one explicit Schur matvec
 * (-1 if S_s are not formed), averaged over `reps` launches, in ms.       */
int  fwh_time_kernels(fwh_ctx *ctx, int reps, double *ms_solve,
                      double *ms_schur, double *ms_smv);

/* Subdomain arrays for inspection.  what: 0 = assembled lower-triangle
 * values (2*nnz_s doubles, CSR order), 1 = CSR row pointer (n_s+1 ints),
 * 2 = CSR columns (nnz_s ints), 3 = load f (2*n_s doubles),
 * 4 = skyline idx1 (n_s ints), 5 = skyline idx2 (n_s ints),
 * 6 = factor (2*nlen_s doubles; L strictly lower, 1/d on the diagonal,
 * advlas_skyline.c:123-125), 7 = sizes {n_s, nnz_s, nlen_s, ninbc_s} ints */
int  fwh_get_sub(const fwh_ctx *ctx, int s, int what, void *out);

/* Element kernel alone (tet_MakeElement_EM_E, make_elem_em.c:14-400) on a
 * batch: xyz [nel*12], eps/sigma [nel], pml [nel*6] or NULL, src [nel],
 * joe [nel*24] (4 vertices x 3 complex).  ae: [nel*72] (6x6 complex,
 * row-major), be: [nel*12].                                               */
int  fwh_element_batch(fwh_ctx *ctx, int nel, const double *xyz,
                       const double *eps, const double *sigma, double omega,
                       const double *pml, const int *src, const double *joe,
                       double *ae, double *be);

/* fwh_element_batch without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate (serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as fwh_element_batch, at
 * most FWH_STATIC_MAX_ELEM elements per call (more: FWH_E_LIMIT).  Its
 * device buffers are one fixed array per GPU (about 6 MB, held until the
 * process ends), shared by every call on that GPU: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the
 * caller.  Synchronous: the results are in ae/be on return.               */
#define FWH_STATIC_MAX_ELEM 4096
int  fwh_element_batch_static(fwh_ctx *ctx, int nel, const double *xyz,
                              const double *eps, const double *sigma,
                              double omega, const double *pml, const int *src,
                              const double *joe, double *ae, double *be);

/* Subdomain COCG with IC(0) (mymt_CS_AIJ_Solver_COCG,
 * lib/mymt_aij_solver/aij_solver_cg.h:9-44,113-355 with
 * mymt_CS_AIJ_PC_ICC_make/set, aij_pc_func.h:36-167) on every subdomain's
 * Dirichlet matrix with load b (2*n_dof doubles, concatenated), zero
 * initial guess.  x: 2*n_dof doubles; status: per subdomain the reference
 * return value (>=0 iterations, <0 error).                                */
int  fwh_cocg_icc_all(fwh_ctx *ctx, const double *b, double *x, int *status,
                      double eps, double ic_shift);
/* One system in the reference's AIJ_MAT storage (lib/std/types.h:44-61:
 * ai[nf+1], lower-triangle columns aij with the diagonal last, values vij,
 * 2 doubles each) by the same GPU COCG: mymt_CS_AIJ_Solver_COCG
 * (aij_solver_cg.h:9-44, 113-355) with pc = SOLVER_PC_ICC (3, IC(0) with
 * acceleration ic_shift, aij_pc_func.h:36-167) or SOLVER_PC_NONE (0); x is
 * the initial guess in, the solution out (2*nf doubles); *status the
 * reference's return value (>= 0 iterations, SOLVER_STATUS_* < 0).  At most
 * 2,048 unknowns (the kernel's vectors live in LDS).                      */
int  fwh_aij_cocg(fwh_ctx *ctx, int nf, const int *ai, const int *aij, const double *vij,
                  const double *b, double *x, int pc, double eps, double div, double ic_shift,
                  int *status);
/* The same with the IC(0) factor in the reference's AIJ_PC layout
 * (aij_pc_func.h:52-58: the factor L in the matrix's pattern, ai[nf]
 * entries, then d = 1/diag(L), nf entries; 2 doubles each): make = 1
 * computes it on the GPU and, when pcm is not null, writes it to pcm, as
 * the PC_ICC make step of mymt_CS_AIJ_Solver_COCG leaves it in
 * aij_mat->pc_c; make = 0 takes it from pcm and skips the make step, as
 * mymt_CS_AIJ_Solver_COCG_nmpcm (aij_solver_cg.h:68-92) reuses it.
 * pcm is not read or written with pc = SOLVER_PC_NONE.                   */
int  fwh_aij_cocg_pc(fwh_ctx *ctx, int nf, const int *ai, const int *aij, const double *vij,
                     const double *b, double *x, int pc, double eps, double div, double ic_shift,
                     double *pcm, int make, int *status);

#ifdef __cplusplus
}
#endif

#endif /* FWHIP_H */
