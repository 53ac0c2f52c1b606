/*
 * refdrv.c - test driver for the reference-shaped call surface of
 * libfwhip (include/fwh_ref.h).  TEST INFRASTRUCTURE.
 *
 * Builds the reference's own structs from a flat model description, the
 * way hddm_Init leaves them for hddm_solver_CS (lib/hddm/make_PartMesh.c:
 * 58-117, lib/std/types.h): per part a PartMesh with its MTRL tables (element
 * flags -> Permittivity / Conductor / Antenna entries), OPinfo lists and
 * interface counts; per subdomain a DomMesh with nop / flag / crd / nmp,
 * dbc_Axn0 = [essential edges | interface edges], Inbc entries and one
 * ANTENNA per antenna definition (enf: local vertex -> antenna node, Jo /
 * Jor per node).  Then calls hddm_solver_CS_hip_status, or runs the
 * reference's driver loop over set_hddm_func_fullwave_hip, and copies
 * DomMesh.x_c and DomMesh.rvalue back out.  Loaded with global symbols:
 * it defines the reference's hddm_sw, which the hook table sets.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/fwh_ref.h"

typedef struct {
  int nflag;            /* Permittivity entries                            */
  const int *flag_i;
  const double *flag_nu;
  int nR;               /* Conductor entries                               */
  const int *R_i;
  const double *R_sigma;
  int nant;             /* Antenna entries (their element flags)           */
  const int *ant_flag;
  double omega;
} refdrv_mtrl;

static void free_parts(PartMesh *pm, int npart) {
  for (int p = 0; p < npart; p++) {
    for (int k = 0; k < pm[p].ndomain; k++) {
      DomMesh *dm = &pm[p].dm[k];
      free(dm->nop); free(dm->flag); free(dm->crd); free(dm->inbc); free(dm->dbc_Axn0.ibc);
      for (int a = 0; a < pm[p].mtrl.nantenna; a++) {
        free(dm->antenna[a].enf); free(dm->antenna[a].Jo); free(dm->antenna[a].Jor);
      }
      free(dm->antenna);
      free(dm->x_c); free(dm->rvalue.el_B_c); free(dm->rvalue.el_E_c); free(dm->rvalue.el_Je_c);
    }
    for (int q = 0; q < pm[p].npart; q++) free(pm[p].op[q].ibid);
    free(pm[p].op);
    free(pm[p].dm);
    free(pm[p].mtrl.antenna_def);
  }
  free(pm);
}

/* the reference-shaped parts of a flat model; flag: per element material flag */
static PartMesh *make_parts(const fwh_mesh_desc *f, const int *flag, const refdrv_mtrl *mt) {
  const int np = f->npart;
  PartMesh *pm = (PartMesh *)calloc(np, sizeof(PartMesh));
  for (int p = 0; p < np; p++) {
    PartMesh *P = &pm[p];
    P->partid = p;
    P->npart = np;
    P->infree = f->infree[p];
    P->outfree = f->outfree[p];
    P->tfree = f->infree[p] + f->outfree[p];
    P->ap_elem = 4;
    P->nd_elem = 10;
    P->dimension = 3;
    P->mtrl.nflag = mt->nflag;
    P->mtrl.flag_i = (int *)mt->flag_i;
    P->mtrl.flag_nu = (double *)mt->flag_nu;
    P->mtrl.nR = mt->nR;
    P->mtrl.R_i = (int *)mt->R_i;
    P->mtrl.R_sigma = (double *)mt->R_sigma;
    P->mtrl.nantenna = mt->nant;
    P->mtrl.antenna_def = (MTRL_ANTENNA_DEFINITION *)calloc(mt->nant ? mt->nant : 1, sizeof(MTRL_ANTENNA_DEFINITION));
    for (int a = 0; a < mt->nant; a++) P->mtrl.antenna_def[a].flag = mt->ant_flag[a];
    P->mtrl.nantenna_omega = 1;
    P->mtrl.antenna_omega = mt->omega;
    P->op = (OPinfo *)calloc(np, sizeof(OPinfo));
    for (int q = 0; q < np; q++) {
      const int o0 = f->op_off[p * np + q], o1 = f->op_off[p * np + q + 1];
      P->op[q].n_mofree = o1 - o0;
      P->op[q].ibid = (int *)malloc((o1 - o0 + 1) * sizeof(int));
      memcpy(P->op[q].ibid, f->op_ibid + o0, (o1 - o0) * sizeof(int));
    }
    for (int s = 0; s < f->nsub; s++) P->ndomain += f->sub_part[s] == p;
    P->dm = (DomMesh *)calloc(P->ndomain ? P->ndomain : 1, sizeof(DomMesh));
  }
  int *next = (int *)calloc(np, sizeof(int));
  for (int s = 0; s < f->nsub; s++) {
    const int p = f->sub_part[s];
    DomMesh *dm = &pm[p].dm[next[p]];
    dm->partid = p;
    dm->domid = next[p]++;
    dm->npart = np;
    dm->ndomain = pm[p].ndomain;
    const int e0 = f->el_off[s], ne = f->el_off[s + 1] - e0;
    const int a0 = f->ap_off[s], na = f->ap_off[s + 1] - a0;
    dm->nel = ne;
    dm->nap = na;
    dm->nmp = f->nmp[s];
    dm->nnd = na + f->nmp[s];
    dm->ap_elem = 4;
    dm->nd_elem = 10;
    dm->dimension = 3;
    strcpy(dm->elem_type, "3DQuadraticTetrahedron");
    dm->nop = (int *)malloc((10 * (size_t)ne + 1) * sizeof(int));
    memcpy(dm->nop, f->nop + 10 * (size_t)e0, 10 * (size_t)ne * sizeof(int));
    dm->flag = (int *)malloc((ne + 1) * sizeof(int));
    memcpy(dm->flag, flag + e0, ne * sizeof(int));
    dm->crd = (double *)malloc((3 * (size_t)na + 1) * sizeof(double));
    memcpy(dm->crd, f->crd + 3 * (size_t)a0, 3 * (size_t)na * sizeof(double));
    const int b0 = f->bc_off[s], nb = f->bc_off[s + 1] - b0;
    const int i0 = f->ib_off[s], ni = f->ib_off[s + 1] - i0;
    dm->ninbc = ni;
    dm->inbc = (Inbc *)calloc(ni ? ni : 1, sizeof(Inbc));
    dm->dbc_Axn0.need = 0;
    dm->dbc_Axn0.nbc = nb + ni;
    dm->dbc_Axn0.ibc = (int *)malloc((nb + ni + 1) * sizeof(int));
    memcpy(dm->dbc_Axn0.ibc, f->ibc + b0, nb * sizeof(int));
    dm->nbc_A = nb;
    for (int i = 0; i < ni; i++) {
      dm->inbc[i].nfid = f->ib_nfid[i0 + i];
      dm->inbc[i].ln_part = f->ib_ln[i0 + i];
      dm->inbc[i].snid = -1;
      dm->inbc[i].charge_part = -1;
      dm->dbc_Axn0.ibc[nb + i] = f->ib_nfid[i0 + i];
    }
    /* the current source: Joe = -(Jo + i Jor) per vertex, one antenna node per vertex */
    dm->antenna = (ANTENNA *)calloc(mt->nant ? mt->nant : 1, sizeof(ANTENNA));
    for (int a = 0; a < mt->nant; a++) {
      ANTENNA *an = &dm->antenna[a];
      an->antenna = a;
      an->nf = na;
      an->enf = (int *)malloc((na + 1) * sizeof(int));
      an->Jo = (double *)calloc(3 * (size_t)na + 1, sizeof(double));
      an->Jor = (double *)calloc(3 * (size_t)na + 1, sizeof(double));
      for (int v = 0; v < na; v++) {
        an->enf[v] = v;
        for (int j = 0; j < 3; j++) {
          an->Jo[3 * v + j] = f->joe ? -f->joe[6 * ((size_t)a0 + v) + 2 * j] : 0.0;
          an->Jor[3 * v + j] = f->joe ? -f->joe[6 * ((size_t)a0 + v) + 2 * j + 1] : 0.0;
        }
      }
    }
  }
  free(next);
  return pm;
}

/* the flat description hddm_solver_CS_hip would load from those structs */
int refdrv_desc(const fwh_mesh_desc *f, const int *flag, const refdrv_mtrl *mt, fwh_mesh_desc *out) {
  PartMesh *pm = make_parts(f, flag, mt);
  const int rc = fwh_ref_desc_from_partmesh(pm, out);
  free_parts(pm, f->npart);
  return rc;
}

void refdrv_desc_free(fwh_mesh_desc *d) { fwh_ref_desc_free(d); }

/* hddm_solver_CS_hip on those structs: x (2 n_dof doubles), B/E/J (6 per
 * element) from DomMesh.x_c / rvalue; the residual log goes to logpath   */
int refdrv_solve(const fwh_mesh_desc *f, const int *flag, const refdrv_mtrl *mt, double eps, int max_loop,
                 const char *logpath, double *x, double *B, double *E, double *J, fwh_stats *st) {
  PartMesh *pm = make_parts(f, flag, mt);
  OPTIONS options;
  memset(&options, 0, sizeof(options));
  options.hddm_opt.hddm_eps = eps;
  options.hddm_opt.hddm_max_loop = max_loop;
  options.hddm_opt.hddm_pc = 1;
  options.hddm_opt.hddm_log = logpath ? 0 : -1;
  options.hddm_opt.fp = logpath ? fopen(logpath, "w") : NULL;
  const int rc = hddm_solver_CS_hip_status(pm, options, st);
  if (options.hddm_opt.fp) fclose(options.hddm_opt.fp);
  if (rc == 0 && max_loop > 0) {
    size_t xo = 0, eo = 0;
    for (int p = 0; p < f->npart; p++)
      for (int k = 0; k < pm[p].ndomain; k++) {
        const DomMesh *dm = &pm[p].dm[k];
        memcpy(x + 2 * xo, dm->x_c, dm->nmp * sizeof(Complex));
        memcpy(B + 6 * eo, dm->rvalue.el_B_c, 3 * (size_t)dm->nel * sizeof(Complex));
        memcpy(E + 6 * eo, dm->rvalue.el_E_c, 3 * (size_t)dm->nel * sizeof(Complex));
        memcpy(J + 6 * eo, dm->rvalue.el_Je_c, 3 * (size_t)dm->nel * sizeof(Complex));
        xo += dm->nmp;
        eo += dm->nel;
      }
  }
  free_parts(pm, f->npart);
  return rc;
}

/* the reference's solver state, as lib/hddm/make_PartMesh.c:16 defines it
 * (the hook table's iteration hooks set its status) */
HDDM_SW hddm_sw;

#define SW_FIRST 0 /* lib/hddm/libhddm.h:37-39 */
#define SW_LAST 2

static void free_vec(HDDM_VEC *v) {
  free(v->p_c); free(v->q_c); free(v->r_c); free(v->z_c); free(v->u_c); free(v->pc_c);
  memset(v, 0, sizeof(*v));
}

static void free_dom_vec(HDDM_DOM_VEC *v) {
  free(v->vec_c); free(v->pc_c);
  memset(v, 0, sizeof(*v));
}

/* the reference's s-mode parent loop over a hook table, restated for the
 * test (lib/hddm/hddm_solver.h:59-183 with the loops of
 * lib/hddm/hddm_solver_s.h:7-140): every part of pm in one process; the
 * interface count is logged as the whole-call entry point logs it.
 * Returns the number of main-loop iterations.                            */
static int parent_loop(PartMesh *pm, int npart, OPTIONS options, void (*set_fn)(HDDM_FUNC *, OPTIONS)) {
  HDDM_FUNC f;
  HDDM_Opt ho;
  SolvOpt so;
  memset(&f, 0, sizeof(f));
  memset(&ho, 0, sizeof(ho));
  memset(&so, 0, sizeof(so));
  HDDM_VEC *vec = NULL;
  hddm_sw.status = SW_FIRST;
  hddm_sw.loop = 0;
  set_fn(&f, options);
  f.general_func_1st(pm, vec, options, ho, so);
  f.set_HDDM_Opt(&ho, options);
  f.set_SolvOpt(&so, options);
  vec = (HDDM_VEC *)calloc(npart, sizeof(HDDM_VEC));
  for (int p = 0; p < npart; p++) f.make_HDDM_VEC(&vec[p], pm[p], ho);
  if (ho.hddm_log == 0 && ho.fp) fprintf(ho.fp, "# DOF on interface %d\n", vec[0].total_free);
  f.general_func_2nd(pm, vec, options, ho, so);
  int iters = 0;
  if (ho.hddm_max_loop == 0) {
    fprintf(stdout, "HDDM Solver was not executed!!!\n");
  } else {
    /* LOOP_1ST */
    for (int p = 0; p < npart; p++)
      for (int k = 0; k < pm[p].ndomain; k++) {
        HDDM_DOM_VEC dv;
        memset(&dv, 0, sizeof(dv));
        f.init_bc(&pm[p].dm[k]);
        f.make_TMP_AIJ(&pm[p].dm[k], pm[p].mtrl);
        f.solve_dom_1st(&pm[p].dm[k], ho, so, f.dom_solver, &dv, f.makePC, so.fp);
        f.COM_res_sp_1st(&vec[p], dv, pm[p].dm[k], ho);
        free_dom_vec(&dv);
      }
    f.iteration_s_1st(pm, vec, ho, f.setPC, f.gatherPC_s);
    hddm_sw.loop++;
    f.general_func_3rd(pm, vec, options, ho, so);
    /* main loop */
    while (hddm_sw.status != SW_LAST) {
      for (int p = 0; p < npart; p++)
        for (int k = 0; k < pm[p].ndomain; k++) {
          HDDM_DOM_VEC dv;
          memset(&dv, 0, sizeof(dv));
          f.renewal_bc_sp(&pm[p].dm[k], vec[p]);
          f.solve_dom(&pm[p].dm[k], so, f.dom_solver_nmpcm, &dv, so.fp);
          f.COM_res_sp(&vec[p], dv, pm[p].dm[k], ho);
          free_dom_vec(&dv);
        }
      f.iteration_s(pm, vec, ho, f.setPC);
      iters++;
      if (hddm_sw.loop == ho.hddm_max_loop) hddm_sw.status = SW_LAST;
      hddm_sw.loop++;
    }
    f.general_func_4th(pm, vec, options, ho, so);
    /* LOOP_LAST */
    for (int p = 0; p < npart; p++) {
      for (int k = 0; k < pm[p].ndomain; k++) {
        HDDM_DOM_VEC dv;
        memset(&dv, 0, sizeof(dv));
        f.renewal_bc_sp(&pm[p].dm[k], vec[p]);
        f.solve_dom_last(&pm[p].dm[k], so, f.dom_solver_nmpcm, &dv, so.fp);
        f.COM_res_sp(&vec[p], dv, pm[p].dm[k], ho);
        free_dom_vec(&dv);
      }
      f.get_result_sp(&pm[p]);
      f.post_sp(&pm[p]);
    }
    f.general_func_5th(pm, vec, options, ho, so);
    for (int p = 0; p < npart; p++) f.out_result(hddm_sw.t_loop, pm[p], options);
  }
  for (int p = 0; p < npart; p++) free_vec(&vec[p]);
  free(vec);
  return iters;
}

/* the hooks the s-mode parent loop calls (hddm_solver.h:59-183,
 * hddm_solver_s.h:7-140), as set_hddm_func_fullwave_hip fills them alone
 * (no GPU call): the number left NULL, and set_HDDM_Opt / set_SolvOpt
 * copying the options (1 when they do)                                    */
int refdrv_hooks_null(int *opts_copied) {
  HDDM_FUNC f;
  OPTIONS options;
  memset(&f, 0, sizeof(f));
  memset(&options, 0, sizeof(options));
  options.hddm_opt.hddm_eps = 1e-8;
  options.hddm_opt.hddm_max_loop = 123;
  options.solvopt.solver_eps = 1e-9;
  set_hddm_func_fullwave_hip(&f, options);
  const void *used[] = {
      (const void *)f.set_HDDM_Opt, (const void *)f.set_SolvOpt, (const void *)f.make_HDDM_VEC,
      (const void *)f.init_bc, (const void *)f.renewal_bc_sp, (const void *)f.make_TMP_AIJ,
      (const void *)f.free_TMP_AIJ, (const void *)f.dom_solver, (const void *)f.dom_solver_nmpcm,
      (const void *)f.makePC, (const void *)f.setPC, (const void *)f.gatherPC_s, (const void *)f.solve_dom_1st,
      (const void *)f.solve_dom, (const void *)f.solve_dom_last, (const void *)f.COM_res_sp_1st,
      (const void *)f.COM_res_sp, (const void *)f.iteration_s_1st, (const void *)f.iteration_s,
      (const void *)f.get_result_sp, (const void *)f.post_sp, (const void *)f.out_result,
      (const void *)f.general_func_1st, (const void *)f.general_func_2nd, (const void *)f.general_func_3rd,
      (const void *)f.general_func_4th, (const void *)f.general_func_5th};
  int nnull = 0;
  for (size_t i = 0; i < sizeof(used) / sizeof(used[0]); i++) nnull += used[i] == NULL;
  HDDM_Opt ho;
  SolvOpt so;
  memset(&ho, 0, sizeof(ho));
  memset(&so, 0, sizeof(so));
  if (f.set_HDDM_Opt) f.set_HDDM_Opt(&ho, options);
  if (f.set_SolvOpt) f.set_SolvOpt(&so, options);
  *opts_copied = ho.hddm_eps == 1e-8 && ho.hddm_max_loop == 123 && so.solver_eps == 1e-9;
  return nnull;
}

/* the same model through the reference's driver loop over
 * set_hddm_func_fullwave_hip: x, B/E/J as refdrv_solve returns them; the
 * main-loop iterations in *iters                                         */
void refdrv_solve_hooks(const fwh_mesh_desc *f, const int *flag, const refdrv_mtrl *mt, double eps, int max_loop,
                        const char *logpath, double *x, double *B, double *E, double *J, int *iters) {
  PartMesh *pm = make_parts(f, flag, mt);
  OPTIONS options;
  memset(&options, 0, sizeof(options));
  options.hddm_opt.hddm_eps = eps;
  options.hddm_opt.hddm_max_loop = max_loop;
  options.hddm_opt.hddm_pc = 1;
  options.hddm_opt.hddm_log = logpath ? 0 : -1;
  options.hddm_opt.fp = logpath ? fopen(logpath, "w") : NULL;
  *iters = parent_loop(pm, f->npart, options, set_hddm_func_fullwave_hip);
  if (options.hddm_opt.fp) fclose(options.hddm_opt.fp);
  if (max_loop > 0) {
    size_t xo = 0, eo = 0;
    for (int p = 0; p < f->npart; p++)
      for (int k = 0; k < pm[p].ndomain; k++) {
        const DomMesh *dm = &pm[p].dm[k];
        memcpy(x + 2 * xo, dm->x_c, dm->nmp * sizeof(Complex));
        memcpy(B + 6 * eo, dm->rvalue.el_B_c, 3 * (size_t)dm->nel * sizeof(Complex));
        memcpy(E + 6 * eo, dm->rvalue.el_E_c, 3 * (size_t)dm->nel * sizeof(Complex));
        memcpy(J + 6 * eo, dm->rvalue.el_Je_c, 3 * (size_t)dm->nel * sizeof(Complex));
        xo += dm->nmp;
        eo += dm->nel;
      }
  }
  free_parts(pm, f->npart);
}

/* advlas_skyline_ldl_decomposite then advlas_skyline_ldl_solve on one
 * DMatrix: mat (2 nlen doubles) is factored in place, work gets 1/d, v is
 * solved in place                                                        */
void refdrv_advlas(int n, int *idx1, int *idx2, double *mat, double *work, double *v) {
  DMatrix dm;
  memset(&dm, 0, sizeof(dm));
  dm.ndim = n;
  dm.nlen = idx2[n - 1] + n;
  dm.idx1 = idx1;
  dm.idx2 = idx2;
  dm.mat_dc = (double _Complex *)mat;
  advlas_skyline_ldl_decomposite(dm, (double _Complex *)work);
  advlas_skyline_ldl_solve(dm, (double _Complex *)v);
}

/* mymt_CS_AIJ_Solver_COCG on an AIJ_MAT and SolvOpt filled as the reference's
 * callers fill them (the Dirichlet matrix of aij_matrix_func.h:140-235,
 * solver options of main.c:40-60): returns the solver's value, x in/out */
int refdrv_mymt_cocg(int nf, int *ai, int *aij, double *vij, double *b, double *x, int pc, double eps,
                     double ic_shift) {
  AIJ_MAT am;
  SolvOpt so;
  memset(&am, 0, sizeof(am));
  memset(&so, 0, sizeof(so));
  am.nf = nf;
  am.ai = ai;
  am.aij = aij;
  am.vij_c = (Complex *)vij;
  am.b_c = (Complex *)b;
  am.x_c = (Complex *)x;
  so.solver_pc = pc;
  so.solver_eps = eps;
  so.solver_div = 1.0e+10; /* DEFAULT_SOLVER_DIV, lib/std/defs.h:157 */
  so.solver_ic_shift = ic_shift;
  so.solver_max_loop = 10 * nf;
  return mymt_CS_AIJ_Solver_COCG(&am, so, NULL);
}

/* the reference's order of the two subdomain solver hooks
 * (HDDM_FUNC.dom_solver, then dom_solver_nmpcm on the same AIJ_MAT with
 * another right-hand side): mymt_CS_AIJ_Solver_COCG from b1 into x1, its
 * preconditioner copied to pcm (pc_size entries, 0 when none was left),
 * then mymt_CS_AIJ_Solver_COCG_nmpcm from b2 into x2; returns the first
 * value, *st2 the second */
int refdrv_mymt_cocg_nmpcm(int nf, int *ai, int *aij, double *vij, double *b1, double *x1, double *b2, double *x2,
                           int pc, double eps, double ic_shift, double *pcm, int *pc_size, int *st2) {
  AIJ_MAT am;
  SolvOpt so;
  memset(&am, 0, sizeof(am));
  memset(&so, 0, sizeof(so));
  am.nf = nf;
  am.ai = ai;
  am.aij = aij;
  am.vij_c = (Complex *)vij;
  am.b_c = (Complex *)b1;
  am.x_c = (Complex *)x1;
  so.solver_pc = pc;
  so.solver_eps = eps;
  so.solver_div = 1.0e+10;
  so.solver_ic_shift = ic_shift;
  so.solver_max_loop = 10 * nf;
  const int st1 = mymt_CS_AIJ_Solver_COCG(&am, so, NULL);
  *pc_size = am.pc_size;
  if (am.pc_size > 0) memcpy(pcm, am.pc_c, sizeof(Complex) * (size_t)am.pc_size);
  am.b_c = (Complex *)b2;
  am.x_c = (Complex *)x2;
  *st2 = mymt_CS_AIJ_Solver_COCG_nmpcm(&am, so, NULL);
  free(am.pc_c); /* mymt_CS_AIJ_Free's Free of the factor */
  return st1;
}
