/*
 * fwh_ref.h - the reference's own call surface for the HDDM full-wave path,
 * backed by libfwhip.so.
 *
 * A maintainer of ADVENTURE_FullWave links libfwhip.so in place of the
 * reference's s-mode solver for this path and calls the same names with the
 * same structs:
 *
 *   hddm_solver_CS(pm, options, set_hddm_func_fullwave)      main.c:105
 *     -> hddm_solver_CS_hip(pm, options)
 *   advlas_skyline_ldl_decomposite(DMatrix, double _Complex *)
 *   advlas_skyline_ldl_solve(DMatrix, double _Complex *)     advlas_skyline.h:11-12
 *     -> the same names, computed on the GPU
 *
 * The structs are the reference's (lib/std/types.h), restated below field
 * for field because their layout is the interface: the library reads
 * PartMesh / DomMesh / MTRL / OPinfo / Inbc / DirichletBC / ANTENNA and
 * writes DomMesh.x_c and DomMesh.rvalue exactly where the reference's own
 * hooks do.  Define FWH_USE_REFERENCE_TYPES and put the reference's
 * lib/std directory on the include path to compile against its types.h
 * instead.  C only (DMatrix holds a C99 double _Complex pointer).
 */
#ifndef FWH_REF_H
#define FWH_REF_H

#include <stdio.h>

#include "fwhip.h"

#ifdef __cplusplus
#error "fwh_ref.h is a C header (the reference's structs use C99 complex types)"
#endif

#ifdef FWH_USE_REFERENCE_TYPES
#include "types.h"
#else

/* ---- lib/std/types.h:15-18 --------------------------------------------- */
typedef struct {
  double re;
  double im;
} Complex;

/* ---- lib/std/types.h:22-61: row-list matrix and CSR ------------------- */
typedef struct {
  int nf;
  int *tmp_ai;          /* entries per row                                  */
  int **tmp_aij;        /* columns per row (j <= i)                         */
  double **tmp_vij_d;
  Complex **tmp_vij_c;
  double *tmp_b_d;
  Complex *tmp_b_c;
  double *tmp_x_d;
  Complex *tmp_x_c;
  int nMargin;
  int *nSize;
  int tmp_pc_size;
  double *tmp_pc_d;
  Complex *tmp_pc_c;
} TMP_AIJ;

typedef struct {
  int nf;
  int *ai;
  int *aij;
  double *vij_d;
  Complex *vij_c;
  double *b_d;
  Complex *b_c;
  double *x_d;
  Complex *x_c;
  int pc_size;
  double *pc_d;
  Complex *pc_c;
} AIJ_MAT;

/* ---- lib/std/types.h:65-92: data switches ------------------------------ */
typedef struct {
  char *data_dir, *fwdata_dir, *fwdata_file, *onedata_dir, *onedata_file;
  char *modeldata_dir, *modeldata_file, *Jometis_dir, *Jometis_file;
  char *Inidata_dir, *Inidata_file, *Inimetis_dir, *Inimetis_file;
  char *outdata_dir, *outdata_file;
  int avsfile, avsfile_micro, vtkfile;
  char *avsfile_dir, *avsfile_file, *vtkfile_dir, *vtkfile_file;
  char *calclog_dir, *calclog_file, *mtrldat_dir, *mtrldat_file;
} DataSW;

/* ---- lib/std/types.h:96-141: solver options ---------------------------- */
typedef struct {
  FILE *fp;
  int solver;
  int solver_dom;
  int solver_pc;
  double solver_eps;
  double solver_div;
  int solver_max_loop;
  double solver_ic_shift;
  int solver_log;
  int coJo_sol;
  int coJo_sol_pc;
  double coJo_sol_eps;
  int coJo_sol_max_loop;
  double coJo_sol_div;
  double coJo_sol_ic_shift;
  int coJo_sol_log;
} SolvOpt;

typedef struct {
  FILE *fp;             /* interface residual log (calc_log)               */
  int hddm_solver;
  int hddm_pc;          /* HDDM_PC_NONE 0 / HDDM_PC_DIAG 1                  */
  double hddm_eps;      /* ||r||inf / ||r0||inf stop                       */
  double hddm_nl_eps;
  double hddm_div;
  double hddm_nl_div;
  int hddm_max_loop;
  int hddm_log;         /* NEED 0 / NO_NEED -1                              */
  int coJo_hddm_sol;
  int coJo_hddm_pc;
  double coJo_hddm_eps;
  int coJo_hddm_max_loop;
  double coJo_hddm_div;
  int coJo_hddm_log;
} HDDM_Opt;

/* ---- lib/std/types.h:143-157: skyline matrix ---------------------------
 * row i holds columns idx1[i]..i at mat_dc[idx2[i] + j]                   */
typedef struct {
  int type;
  int ndim;
  int nlen;
  int *idx1;
  int *idx2;
  double *mat;
  Complex *mat_c;
  double _Complex *mat_dc;
  double *rhs;
  Complex *rhs_c;
} DMatrix;

/* ---- lib/std/types.h:160-206: options ---------------------------------- */
typedef struct {
  int CurrentDensity;
  int MagneticField;
} UNIT;

typedef struct {
  DataSW datasw;
  SolvOpt solvopt;
  HDDM_Opt hddm_opt;
  UNIT unit;
  int work;
  int field;
  int analysis;
  int formulation;
  size_t memlimit;
  int trnFile;
  char *fn_trnFile;
  int tmpFile;
  char *fn_tmpFile;
  int femFile;
  char *fn_femFile;
  double crd_magnification;
  int start_step;
  int time_step;
  int out_interval;
  double delta_t;
  int nl_method;
  int nl_max_loop;
  double nl_eps;
  double nl_div;
  int nl_curve;
  int nl_curve_ip;
  double nl_picard_weight;
} OPTIONS;

/* ---- lib/std/types.h:222-286: materials --------------------------------- */
typedef struct {
  int n;
  double ap[5][3];
  char analysis[128];
  double Jor[3];
  double Joi[3];
  double abc[3][3];
  double d[2][3];
  double tolerance;
} MTRL_ANTENNA_DEF_PP;

typedef struct {
  int n;
  int xyz;
  double ap[4];
  double theta[2];
  double r[2];
  char analysis[128];
  double Jor;
  double Joi;
} MTRL_ANTENNA_DEF_DSC;

typedef struct {
  int flag;             /* element flag of the antenna region              */
  int mode;
  char fname[FILENAME_MAX];
  int ndef;
  int npp;
  int ndsc;
  MTRL_ANTENNA_DEF_PP *pp;
  MTRL_ANTENNA_DEF_DSC *dsc;
} MTRL_ANTENNA_DEFINITION;

typedef struct {
  int nvol;
  int *vol;
  int nbc;
  int *bci;
  int nflag;            /* Permittivity entries: flag_i -> flag_nu (eps)   */
  int *flag_i;
  double *flag_nu;
  int nantenna;
  MTRL_ANTENNA_DEFINITION *antenna_def;
  int nantenna_omega;
  double antenna_omega; /* omega of the analysis                           */
  int nR;               /* Conductor entries: R_i -> R_sigma               */
  int *R_i;
  double *R_sigma;
} MTRL;

/* ---- lib/std/types.h:290-330: boundary data ----------------------------- */
typedef struct {
  int need;
  int nbc;
  int *ibc;
  double *vbc;
  Complex *vbc_c;
} DirichletBC;

typedef struct {
  int snid;
  int nfid;             /* local edge (mid-node) id of the interface DOF    */
  int coordinate;
  int charge_part;
  int ln_chpar;
  int ln_part;          /* index into the part's interface vectors          */
} Inbc;

typedef struct {
  int n_mofree;
  int *ibid;
} OPinfo;

typedef struct {
  int antenna;
  int nf;
  int *enf;             /* local vertex -> antenna node                     */
  int *apid2gnid;
  double *Jo;           /* 3 per antenna node: Re(A) source                 */
  double *Jor;          /* 3 per antenna node: Im(A) source                 */
  double *Ihr;
  double *Ihi;
  DirichletBC dbc;
  TMP_AIJ tmp_aij;
  AIJ_MAT aij_mat;
  int ninbc;
  Inbc *inbc;
} ANTENNA;

typedef struct {
  double *B_d;
  Complex *B_c;
  double *E_d;
  Complex *E_c;
  double *Je_d;
  Complex *Je_c;
  double *el_B_d;
  Complex *el_B_c;
  double *el_E_d;
  Complex *el_E_c;
  double *el_Je_d;
  Complex *el_Je_c;
} ResultValue;

/* ---- lib/std/types.h:417-484: a subdomain ------------------------------ */
typedef struct {
  int ninbc;
  Inbc *inbc;
} DomANTENNA;

typedef struct {
  int partid;
  int domid;
  int npart;
  int ndomain;
  int nel;
  int nnd;
  int nap;              /* vertices                                         */
  int nmp;              /* mid-nodes = edge DOFs                            */
  int nphi;
  int *nop;             /* 10 per element: 4 vertex ids, 6 mid-node ids     */
  int *flag;            /* material flag per element                        */
  double *crd;          /* 3 per vertex                                     */
  int ap_elem;
  int nd_elem;
  char elem_type[128];
  int dimension;
  int ifd_dimension;
  int ninbc;
  Inbc *inbc;           /* interface DOFs, sorted by nfid                   */
  int *snid2pnid;
  int *seid2geid;
  int *apid2phiid;
  int *apid2snid;
  int *mpid2snid;
  TMP_AIJ tmp_aij;
  DMatrix dmat;
  double *x_d;
  Complex *x_c;         /* solution per edge DOF (set by the solve)         */
  double *b_d_mem;
  Complex *b_c_mem;
  DirichletBC dbc_Axn0; /* [essential edges | interface edges]             */
  DirichletBC dbc_Axn0_backup;
  int nbc_A;
  ANTENNA *antenna;     /* one per MTRL antenna definition                  */
  ResultValue rvalue;   /* el_B_c / el_E_c / el_Je_c set by the post step   */
} DomMesh;

/* ---- lib/std/types.h:488-563: a part ------------------------------------ */
typedef struct {
  int nf;
  int yng_npart;
  int *yng_part;
  int *yng_nf_part;
  int old_npart;
  int *old_part;
  int *old_nf_part;
  int charge_part;
  int ln_chpar;
  int ln_part;
  int bc_flag;
} Pinbc;

typedef struct {
  int npart;
  int ndomain;
  int infree;
  int outfree;
  int tfree;
  int total_free;
  OPinfo *op;
  DomANTENNA *dantenna;
} PartANTENNA;

typedef struct {
  int partid;
  int npart;
  int ndomain;
  int nel;
  int nel_all;
  int nnd;
  int nnd_all;
  int nap;
  int nmp;
  double *crd;
  int ap_elem;
  int nd_elem;
  char elem_type[128];
  int dimension;
  int *pnid2gnid;
  int infree;           /* interface DOFs owned by this part                */
  int outfree;          /* copies owned by a lower part (ln_part < outfree) */
  int midfree;
  int tfree;
  int total_free;
  OPinfo *op;           /* [npart]: ln_part lists shared with each part     */
  int ifd_dimension;
  Pinbc *pinbc;
  MTRL mtrl;
  PartANTENNA *pantenna;
  PartANTENNA presult;
  DirichletBC dbc_Axn0;
  DirichletBC dbc_Axn0_backup;
  double *hddm_vec_u_d_mem;
  Complex *hddm_vec_u_c_mem;
  DomMesh *dm;          /* [ndomain]                                        */
  int n_outdom;
  int sum_ninbd;
} PartMesh;

/* ---- lib/std/types.h:566-608: a part's interface vectors --------------- */
typedef struct {
  int infree;
  int outfree;
  int tfree;
  int total_free;
  double *p_d;
  Complex *p_c;
  double *q_d;
  Complex *q_c;
  double *r_d;
  Complex *r_c;
  double *z_d;
  Complex *z_c;
  double *u_d;
  Complex *u_c;
  double r0;
  double rz_d;
  Complex rz_c;
  int pc_size;
  double *pc_d;
  Complex *pc_c;
} HDDM_VEC;

/* ---- lib/std/types.h:611-626: a subdomain's contribution --------------- */
typedef struct {
  int nf;
  double *vec_d;
  Complex *vec_c;
  int pc_size;
  double *pc_d;
  Complex *pc_c;
} HDDM_DOM_VEC;

/* ---- lib/std/types.h:629-730: the hook table --------------------------- */
typedef int (*fwh_ref_dom_solver_fn)(AIJ_MAT *aij_mat, SolvOpt solvopt, FILE *monitor);
typedef int (*fwh_ref_make_pc_fn)(TMP_AIJ tmp_aij, HDDM_DOM_VEC *hddm_dom_vec);
typedef int (*fwh_ref_set_pc_fn)(HDDM_VEC *hddm_vec);
typedef struct {
  void (*set_HDDM_Opt)(HDDM_Opt *hddm_opt, OPTIONS options);
  void (*set_SolvOpt)(SolvOpt *solvopt, OPTIONS options);
  void (*make_HDDM_VEC)(HDDM_VEC *hddm_vec, PartMesh pm, HDDM_Opt hddm_opt);
  void (*init_bc)(DomMesh *dm);
  void (*renewal_bc_sp)(DomMesh *dm, HDDM_VEC hddm_vec);
  void (*renewal_bc_h_send_dom)(DomMesh dm, HDDM_VEC hddm_vec, int compid, int type);
  void (*renewal_bc_h_recv_dom)(DomMesh *dm, int compid, int type);
  void (*make_TMP_AIJ)(DomMesh *dm, MTRL mtrl);
  void (*free_TMP_AIJ)(DomMesh *dm);
  fwh_ref_dom_solver_fn dom_solver;
  fwh_ref_dom_solver_fn dom_solver_nmpcm;
  fwh_ref_make_pc_fn makePC;
  fwh_ref_set_pc_fn setPC;
  void (*gatherPC_s)(OPinfo **oop, HDDM_VEC *hddm_vec);
  void (*gatherPC_ph)(OPinfo *op, HDDM_VEC *hddm_vec);
  void (*solve_dom_1st)(DomMesh *dm, HDDM_Opt hddm_opt, SolvOpt solvopt, fwh_ref_dom_solver_fn dom_solver,
                        HDDM_DOM_VEC *hddm_dom_vec, fwh_ref_make_pc_fn makePC, FILE *fp_log);
  void (*solve_dom)(DomMesh *dm, SolvOpt solvopt, fwh_ref_dom_solver_fn dom_solver_nmpcm,
                    HDDM_DOM_VEC *hddm_dom_vec, FILE *fp_log);
  void (*solve_dom_last)(DomMesh *dm, SolvOpt solvopt, fwh_ref_dom_solver_fn dom_solver_nmpcm,
                         HDDM_DOM_VEC *hddm_dom_vec, FILE *fp_log);
  void (*COM_res_sp_1st)(HDDM_VEC *hddm_vec, HDDM_DOM_VEC hddm_dom_vec, DomMesh dm, HDDM_Opt hddm_opt);
  void (*COM_res_sp)(HDDM_VEC *hddm_vec, HDDM_DOM_VEC hddm_dom_vec, DomMesh dm, HDDM_Opt hddm_opt);
  void (*COM_res_h_recv_1st)(HDDM_VEC *hddm_vec, DomMesh dm, HDDM_Opt hddm_opt, int compid, int type);
  void (*COM_res_h_recv)(HDDM_VEC *hddm_vec, DomMesh dm, HDDM_Opt hddm_opt, int compid, int type);
  void (*iteration_s_1st)(PartMesh *pm, HDDM_VEC *hddm_vec, HDDM_Opt hddm_opt, fwh_ref_set_pc_fn setPC,
                          void (*gatherPC_s)(OPinfo **oop, HDDM_VEC *hddm_vec));
  void (*iteration_s)(PartMesh *pm, HDDM_VEC *hddm_vec, HDDM_Opt hddm_opt, fwh_ref_set_pc_fn setPC);
  void (*iteration_ph_1st)(PartMesh pm, HDDM_VEC *hddm_vec, HDDM_Opt hddm_opt, fwh_ref_set_pc_fn setPC,
                           void (*gatherPC_ph)(OPinfo *op, HDDM_VEC *hddm_vec));
  void (*iteration_ph)(PartMesh pm, HDDM_VEC *hddm_vec, HDDM_Opt hddm_opt, fwh_ref_set_pc_fn setPC);
  void (*get_result_sp)(PartMesh *pm);
  void (*post_sp)(PartMesh *pm);
  void (*post_h_recv_res)(DomMesh *dm, int compid, int type);
  void (*post_h_send_res)(DomMesh *dm, MTRL mtrl, int compid, int type);
  void (*out_result)(int t_loop, PartMesh pm, OPTIONS options);
  void (*general_func_1st)(PartMesh *pm, HDDM_VEC *hddm_vec, OPTIONS options, HDDM_Opt hddm_opt, SolvOpt solvopt);
  void (*general_func_2nd)(PartMesh *pm, HDDM_VEC *hddm_vec, OPTIONS options, HDDM_Opt hddm_opt, SolvOpt solvopt);
  void (*general_func_3rd)(PartMesh *pm, HDDM_VEC *hddm_vec, OPTIONS options, HDDM_Opt hddm_opt, SolvOpt solvopt);
  void (*general_func_4th)(PartMesh *pm, HDDM_VEC *hddm_vec, OPTIONS options, HDDM_Opt hddm_opt, SolvOpt solvopt);
  void (*general_func_5th)(PartMesh *pm, HDDM_VEC *hddm_vec, OPTIONS options, HDDM_Opt hddm_opt, SolvOpt solvopt);
} HDDM_FUNC;

/* ---- lib/std/types.h:732-741: the solver's state (global hddm_sw,
 * lib/hddm/make_PartMesh.c:16) -------------------------------------------- */
typedef struct {
  int status;
  int loop;
  int ndomain;
  int antenna;
  int Jo;
  int nl_loop;
  int t_loop;
} HDDM_SW;

#endif /* FWH_USE_REFERENCE_TYPES */

/* ---- entry points ------------------------------------------------------ */

/* hddm_solver_CS(pm, options, set_hddm_func_fullwave) in s-mode
 * (lib/hddm/hddm_solver_cs.h:15-17, lib/hddm/hddm_solver.h:25-183 with the
 * full-wave hooks of main/fullwave/fullwave.c:23-146), called at
 * main/fullwave/main.c:105, on the GPU.  pm: the npart parts of the model
 * (pm[0].npart of them) as hddm_Init leaves them
 * (lib/hddm/make_PartMesh.c:58-117).  Reads, per subdomain, nop / flag /
 * crd / nmp / dbc_Axn0 / inbc / antenna and, per part, mtrl / op / infree /
 * outfree; the source comes from ./MODEL/volt.dat as in the reference
 * (lib/fem/make_elem_em.c:74-96).  On return, as after the reference's
 * LOOP_LAST, get_result_sp and post_sp: dm->x_c (nmp Complex, malloc) and
 * dm->rvalue.el_B_c / el_E_c / el_Je_c (3 Complex per element, malloc) of
 * every subdomain.  With options.hddm_opt.hddm_log == NEED (0), the
 * interface count and the residual history go to options.hddm_opt.fp (the
 * calc_log the caller opened, as HDDM_set_HDDM_Opt_log does,
 * lib/hddm/set_hddm_func.c:37-52) in the reference's format.  The result files
 * are left to the caller's out_result (write_outdata_file_HDDM_FW,
 * lib/std/fileio.c:2930-3224), which reads those same fields.
 * Errors print a message and exit, as emessage does
 * (lib/std/errormessage.h:13); hddm_solver_CS_hip_status returns them
 * instead (0 or an FWH_E_* code, message in fwh_last_error()).            */
void hddm_solver_CS_hip(PartMesh *pm, OPTIONS options);
int  hddm_solver_CS_hip_status(PartMesh *pm, OPTIONS options, fwh_stats *stats);

/* set_hddm_func_fullwave (main/fullwave/fullwave.h:17, fullwave.c:23-146)
 * for the reference's own driver: hddm_solver_CS(pm, options,
 * set_hddm_func_fullwave_hip) runs the reference's parent loop
 * (lib/hddm/hddm_solver.h:59-183, hddm_solver_s.h:7-140) over these hooks,
 * which batch the work on the GPU: general_func_1st loads the whole model
 * (as hddm_solver_CS_hip does), the per-subdomain hooks of each loop
 * (init_bc, make_TMP_AIJ, renewal_bc_sp, solve_dom_1st / solve_dom /
 * solve_dom_last, COM_res_sp_1st / COM_res_sp) only queue, iteration_s_1st
 * runs the first step and iteration_s the interface iterations (ahead, in
 * hipGraphs of 16, as many as the loop asks for), setting hddm_sw.status to
 * __HDDM_SW_LAST__ (2) at the loop where the GPU's stop test fired, so the
 * loop runs the reference's iteration count; get_result_sp and post_sp fill
 * dm->x_c and dm->rvalue.el_B_c / el_E_c / el_Je_c of the part, and
 * general_func_5th releases the GPU.  make_HDDM_VEC allocates the part's
 * host vectors (they stay zero: the iterates live on the GPU).  The
 * subdomain solver hooks are mymt_CS_AIJ_Solver_COCG / _nmpcm of this
 * library.  Linked into the reference, the hooks this table does not
 * replace (set_HDDM_Opt, set_SolvOpt, out_result, free_TMP_AIJ and the
 * p/h-mode ones) are the reference's own, taken from its
 * set_hddm_func_fullwave; alone, set_HDDM_Opt / set_SolvOpt copy
 * options.hddm_opt / options.solvopt and the rest do nothing.  hddm_sw is
 * the reference's global (found as a symbol of the process).  One model
 * at a time per process; errors print a message and exit, as emessage
 * does.                                                                   */
void set_hddm_func_fullwave_hip(HDDM_FUNC *hddm_func, OPTIONS options);

/* The flat description libfwhip loads, built from the reference's structs
 * (no reference counterpart: the translation hddm_solver_CS_hip uses).
 * Arrays are malloc'd into *out and released by fwh_ref_desc_free.        */
int  fwh_ref_desc_from_partmesh(const PartMesh *pm, fwh_mesh_desc *out);
void fwh_ref_desc_free(fwh_mesh_desc *d);

/* advlas_skyline_ldl_decomposite (lib/hddm/advlas_skyline.c:79-128):
 * factors dmat.mat_dc (dmat.ndim rows, profile idx1/idx2) in place into L
 * (strictly lower) and 1/d (on the diagonal); work: dmat.ndim entries
 * (receives 1/d, as in the reference).  advlas_skyline_ldl_solve
 * (advlas_skyline.c:134-167): solves with that factor in place.  Both run
 * on the GPU of the process-wide context (device FWH_DEVICE, default 0) and
 * exit with a message on a failure, as the reference's emessage does.     */
void advlas_skyline_ldl_decomposite(DMatrix dmat, double _Complex *work);
void advlas_skyline_ldl_solve(DMatrix dmat, double _Complex *solution);

/* mymt_CS_AIJ_Solver_COCG (lib/mymt_aij_solver/mymt_aij_cs.h:64; the
 * preconditioner switch aij_solver_cg.h:9-44 and the COCG of :113-355) on
 * the GPU of the same process-wide context: the system in aij_mat (ai, aij,
 * vij_c; b_c; x_c the initial guess, overwritten by the solution), stopping
 * on ||r||_2/||b||_2 <= solvopt.solver_eps, diverged at solver_div, at most
 * 10 nf iterations, IC(0) with acceleration solver_ic_shift for
 * SOLVER_PC_ICC or none for SOLVER_PC_NONE; returns the reference's value
 * (>= 0 iterations, SOLVER_STATUS_* < 0; SOLVER_STATUS_UNKNOWN_PC for the
 * other preconditioners).  At most 2,048 unknowns.                         */
int mymt_CS_AIJ_Solver_COCG(AIJ_MAT *aij_mat, SolvOpt solvopt, FILE *monitor);
/* With SOLVER_PC_ICC the IC(0) factor is left in aij_mat->pc_c (pc_size =
 * ai[nf] + nf entries: L in the matrix's pattern, then d = 1/diag(L)), as
 * the reference's make step (aij_pc_func.h:52-58) leaves it; allocated by
 * the reference's counted allocator when the library is linked into it,
 * else by malloc, and released by mymt_CS_AIJ_Free.
 *
 * mymt_CS_AIJ_Solver_COCG_nmpcm (mymt_aij_cs.h:67; aij_solver_cg.h:68-92):
 * the same solve reusing that factor (no make step), as the reference's
 * iteration steps call it through HDDM_FUNC.dom_solver_nmpcm.              */
int mymt_CS_AIJ_Solver_COCG_nmpcm(AIJ_MAT *aij_mat, SolvOpt solvopt, FILE *monitor);

#endif /* FWH_REF_H */
