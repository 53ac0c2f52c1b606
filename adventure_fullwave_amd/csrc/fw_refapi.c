/*
 * fw_refapi.c - the reference's own call surface for the HDDM full-wave
 * path (include/fwh_ref.h), on top of libfwhip's C ABI (include/fwhip.h).
 *
 * hddm_solver_CS_hip takes the reference's PartMesh array as hddm_Init
 * leaves it (lib/hddm/make_PartMesh.c:58-117) and does what
 * hddm_solver_CS(pm, options, set_hddm_func_fullwave) does in s-mode
 * (lib/hddm/hddm_solver.h:59-183 with the hooks of
 * main/fullwave/fullwave.c:23-146) up to post_sp: assembly, factorisation,
 * interface COCG, last solve, element post-processing, all on one GPU.
 * The translation from the reference's structs to libfwhip's flat
 * description (fwh_ref_desc_from_partmesh) resolves per element what the
 * reference's element kernel looks up per call (lib/fem/make_elem_em.c:
 * 74-115: volt.dat, permittivity, conductivity, antenna).
 *
 * advlas_skyline_ldl_decomposite / _solve keep the reference's prototypes
 * (lib/hddm/advlas_skyline.h:11-12) and run on the GPU of a process-wide
 * context.
 */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fwh_ref.h"

#define REF_NU 7.957747e+05  /* 1/mu0 as fixed in make_elem_em.c:37 and fullwave.c:670-675 */
#define REF_NEED 0           /* NEED, lib/std/defs.h:111 */
#define REF_VOLT "./MODEL/volt.dat"
#define REF_SOLVER_PC_NONE 0             /* lib/std/defs.h:137 */
#define REF_SOLVER_PC_ICC 3              /* lib/std/defs.h:140 */
#define REF_SOLVER_STATUS_UNKNOWN_PC (-4) /* lib/std/defs.h:167 */

/* records the message where fwh_last_error() reports it (fw_host.hip) */
int fwhi_fail(int code, const char *msg);

static int ref_fail(int code, const char *msg) { return fwhi_fail(code, msg); }

static void *zalloc(size_t n, size_t size) { return calloc(n ? n : 1, size); }

void fwh_ref_desc_free(fwh_mesh_desc *d) {
  if (!d) return;
  free((void *)d->sub_part); free((void *)d->el_off); free((void *)d->nop);
  free((void *)d->eps); free((void *)d->sigma); free((void *)d->src);
  free((void *)d->ap_off); free((void *)d->crd); free((void *)d->joe);
  free((void *)d->nmp); free((void *)d->bc_off); free((void *)d->ibc);
  free((void *)d->ib_off); free((void *)d->ib_nfid); free((void *)d->ib_ln);
  free((void *)d->infree); free((void *)d->outfree); free((void *)d->op_off);
  free((void *)d->op_ibid); free((void *)d->dgv);
  memset(d, 0, sizeof(*d));
}

/* volt.dat (make_elem_em.c:74-96, :333-347): dg, then per gap box the
 * lines "x0 x1", "y0 y1", "z0 z1", "dV".  Returns dg (>= 0) and the boxes
 * (6 bounds + dV each, malloc'd), or -1 when the file cannot be read.      */
static int read_volt(double **boxes) {
  char buf[256];
  int dg = 0;
  *boxes = NULL;
  FILE *fp = fopen(REF_VOLT, "r");
  if (!fp) return -1;
  if (!fgets(buf, sizeof(buf), fp) || sscanf(buf, "%d", &dg) != 1 || dg < 0) {
    fclose(fp);
    return -1;
  }
  if (dg > 0) {
    double *b = (double *)zalloc((size_t)dg * 7, sizeof(double));
    for (int i = 0; i < dg; i++) {
      double *q = b + 7 * i;
      int ok = fgets(buf, sizeof(buf), fp) && sscanf(buf, "%lf %lf", &q[0], &q[1]) == 2;
      ok = ok && fgets(buf, sizeof(buf), fp) && sscanf(buf, "%lf %lf", &q[2], &q[3]) == 2;
      ok = ok && fgets(buf, sizeof(buf), fp) && sscanf(buf, "%lf %lf", &q[4], &q[5]) == 2;
      ok = ok && fgets(buf, sizeof(buf), fp) && sscanf(buf, "%lf", &q[6]) == 1;
      if (!ok) {
        free(b);
        fclose(fp);
        return -1;
      }
    }
    *boxes = b;
  }
  fclose(fp);
  return dg;
}

/* the reference's per-element material lookups (make_elem_em.c:98-115):
 * the first matching entry wins, no match leaves eps = 0, sigma = 0, no
 * antenna                                                                */
static double lookup_eps(const MTRL *mt, int flg) {
  for (int i = 0; i < mt->nflag; i++)
    if (flg == mt->flag_i[i]) return mt->flag_nu[i];
  return 0.0;
}
static double lookup_sigma(const MTRL *mt, int flg) {
  for (int i = 0; i < mt->nR; i++)
    if (flg == mt->R_i[i]) return mt->R_sigma[i];
  return 0.0;
}
static int lookup_antenna(const MTRL *mt, int flg) {
  for (int i = 0; i < mt->nantenna; i++)
    if (flg == mt->antenna_def[i].flag) return i;
  return -1;
}
static int has_flag(const MTRL *mt, int flg) {
  for (int i = 0; i < mt->nflag; i++)
    if (flg == mt->flag_i[i]) return 1;
  return 0;
}

int fwh_ref_desc_from_partmesh(const PartMesh *pm, fwh_mesh_desc *d) {
  if (!pm || !d) return ref_fail(FWH_E_ARG, "null argument");
  memset(d, 0, sizeof(*d));
  const int npart = pm[0].npart;
  if (npart <= 0) return ref_fail(FWH_E_ARG, "PartMesh.npart <= 0");
  long nsub = 0, nel = 0, nap = 0, nbc = 0, nib = 0, nop_tot = 0;
  for (int p = 0; p < npart; p++) {
    if (pm[p].npart != npart) return ref_fail(FWH_E_ARG, "parts disagree on npart");
    if (pm[p].ndomain < 0 || (pm[p].ndomain > 0 && !pm[p].dm)) return ref_fail(FWH_E_ARG, "part without subdomains");
    for (int s = 0; s < pm[p].ndomain; s++) {
      const DomMesh *dm = &pm[p].dm[s];
      const int ne = dm->dbc_Axn0.nbc - dm->ninbc;
      if (dm->nel < 0 || dm->nap < 0 || dm->nmp < 0 || ne < 0)
        return ref_fail(FWH_E_ARG, "inconsistent subdomain sizes (nel, nap, nmp, dbc_Axn0.nbc - ninbc)");
      nsub++;
      nel += dm->nel;
      nap += dm->nap;
      nbc += ne;
      nib += dm->ninbc;
    }
    for (int q = 0; q < npart; q++) nop_tot += (pm[p].op && q != p) ? pm[p].op[q].n_mofree : 0;
  }
  int *sub_part = (int *)zalloc(nsub, sizeof(int));
  int *el_off = (int *)zalloc(nsub + 1, sizeof(int));
  int *nop = (int *)zalloc(nel * 10, sizeof(int));
  double *eps = (double *)zalloc(nel, sizeof(double));
  double *sigma = (double *)zalloc(nel, sizeof(double));
  int *src = (int *)zalloc(nel, sizeof(int));
  int *ap_off = (int *)zalloc(nsub + 1, sizeof(int));
  double *crd = (double *)zalloc(nap * 3, sizeof(double));
  double *joe = (double *)zalloc(nap * 6, sizeof(double));
  char *joe_set = (char *)zalloc(nap, 1);
  int *nmp = (int *)zalloc(nsub, sizeof(int));
  int *bc_off = (int *)zalloc(nsub + 1, sizeof(int));
  int *ibc = (int *)zalloc(nbc, sizeof(int));
  int *ib_off = (int *)zalloc(nsub + 1, sizeof(int));
  int *ib_nfid = (int *)zalloc(nib, sizeof(int));
  int *ib_ln = (int *)zalloc(nib, sizeof(int));
  int *infree = (int *)zalloc(npart, sizeof(int));
  int *outfree = (int *)zalloc(npart, sizeof(int));
  int *op_off = (int *)zalloc((size_t)npart * npart + 1, sizeof(int));
  int *op_ibid = (int *)zalloc(nop_tot, sizeof(int));
  /* hand every array to the description at once, so that one free releases them */
  d->npart = npart;
  d->nsub = (int)nsub;
  d->omega = pm[0].mtrl.antenna_omega;
  d->sub_part = sub_part; d->el_off = el_off; d->nop = nop; d->eps = eps; d->sigma = sigma;
  d->src = src; d->pml = NULL; d->ap_off = ap_off; d->crd = crd; d->joe = joe; d->nmp = nmp;
  d->bc_off = bc_off; d->ibc = ibc; d->ib_off = ib_off; d->ib_nfid = ib_nfid; d->ib_ln = ib_ln;
  d->infree = infree; d->outfree = outfree; d->op_off = op_off; d->op_ibid = op_ibid; d->dgv = NULL;
  if (!sub_part || !el_off || !nop || !eps || !sigma || !src || !ap_off || !crd || !joe || !joe_set || !nmp ||
      !bc_off || !ibc || !ib_off || !ib_nfid || !ib_ln || !infree || !outfree || !op_off || !op_ibid) {
    free(joe_set);
    fwh_ref_desc_free(d);
    return ref_fail(FWH_E_ARG, "out of host memory");
  }
  int rc = FWH_OK;
  const char *why = NULL;
  long s = 0, e0 = 0, a0 = 0, b0 = 0, i0 = 0;
  for (int p = 0; p < npart && !why; p++) {
    const MTRL *mt = &pm[p].mtrl;
    infree[p] = pm[p].infree;
    outfree[p] = pm[p].outfree;
    for (int k = 0; k < pm[p].ndomain && !why; k++, s++) {
      const DomMesh *dm = &pm[p].dm[k];
      sub_part[s] = p;
      nmp[s] = dm->nmp;
      el_off[s] = (int)e0;
      ap_off[s] = (int)a0;
      bc_off[s] = (int)b0;
      ib_off[s] = (int)i0;
      memcpy(crd + 3 * a0, dm->crd, 3 * (size_t)dm->nap * sizeof(double));
      for (int el = 0; el < dm->nel; el++) {
        const int *ne = dm->nop + 10 * (size_t)el;
        for (int i = 0; i < 10; i++) {
          const int lim = i < 4 ? dm->nap : dm->nmp;
          if (ne[i] < 0 || ne[i] >= lim) {
            why = "element node id out of range (vertices < nap, mid-nodes < nmp)";
            break;
          }
          nop[10 * (e0 + el) + i] = ne[i];
        }
        if (why) break;
        const int flg = dm->flag[el];
        eps[e0 + el] = lookup_eps(mt, flg);
        sigma[e0 + el] = lookup_sigma(mt, flg);
        const int nc = lookup_antenna(mt, flg);
        if (nc >= 0) {
          /* Joe = -(Jo + i Jor) per vertex (make_elem_em.c:286-296) */
          const ANTENNA *an = &dm->antenna[nc];
          src[e0 + el] = 1;
          for (int i = 0; i < 4; i++) {
            const long v = a0 + ne[i];
            const int a = an->enf[ne[i]];
            double val[6];
            for (int j = 0; j < 3; j++) {
              val[2 * j] = -an->Jo[3 * a + j];
              val[2 * j + 1] = -an->Jor[3 * a + j];
            }
            if (joe_set[v] && memcmp(joe + 6 * v, val, sizeof(val)) != 0) {
              why = "a vertex lies in two antenna regions with different currents";
              break;
            }
            memcpy(joe + 6 * v, val, sizeof(val));
            joe_set[v] = 1;
          }
        }
      }
      const int ne_bc = dm->dbc_Axn0.nbc - dm->ninbc;
      for (int i = 0; i < ne_bc && !why; i++) {
        const int b = dm->dbc_Axn0.ibc[i];
        if (b < 0 || b >= dm->nmp) why = "essential boundary DOF out of range";
        ibc[b0 + i] = b;
      }
      for (int i = 0; i < dm->ninbc && !why; i++) {
        ib_nfid[i0 + i] = dm->inbc[i].nfid;
        ib_ln[i0 + i] = dm->inbc[i].ln_part;
      }
      e0 += dm->nel;
      a0 += dm->nap;
      b0 += ne_bc;
      i0 += dm->ninbc;
    }
  }
  el_off[nsub] = (int)e0;
  ap_off[nsub] = (int)a0;
  bc_off[nsub] = (int)b0;
  ib_off[nsub] = (int)i0;
  {
    long o = 0;
    for (int p = 0; p < npart; p++)
      for (int q = 0; q < npart; q++) {
        op_off[p * npart + q] = (int)o;
        if (q == p || !pm[p].op) continue;
        const OPinfo *op = &pm[p].op[q];
        memcpy(op_ibid + o, op->ibid, (size_t)op->n_mofree * sizeof(int));
        o += op->n_mofree;
      }
    op_off[npart * npart] = (int)o;
  }
  free(joe_set);
  if (!why) {
    double *boxes = NULL;
    const int dg = read_volt(&boxes);
    if (dg < 0) {
      why = "cannot read " REF_VOLT;
    } else if (dg > 0) {
      /* per vertex the dV of the last gap box containing it (bounds
       * included), 0 outside every box (make_elem_em.c:350-361)            */
      double *dgv = (double *)zalloc(nap, sizeof(double));
      for (long v = 0; v < nap; v++) {
        const double *x = crd + 3 * v;
        for (int i = 0; i < dg; i++) {
          const double *q = boxes + 7 * i;
          if (q[0] <= x[0] && x[0] <= q[1] && q[2] <= x[1] && x[1] <= q[3] && q[4] <= x[2] && x[2] <= q[5])
            dgv[v] = q[6];
        }
      }
      d->dgv = dgv;
    }
    free(boxes);
  }
  if (why) {
    fwh_ref_desc_free(d);
    rc = ref_fail(FWH_E_ARG, why);
  }
  return rc;
}

static int device_from_env(void) {
  const char *e = getenv("FWH_DEVICE");
  return e ? atoi(e) : 0;
}

int hddm_solver_CS_hip_status(PartMesh *pm, OPTIONS options, fwh_stats *stats) {
  fwh_mesh_desc d;
  int rc = fwh_ref_desc_from_partmesh(pm, &d);
  if (rc) return rc;
  const HDDM_Opt *ho = &options.hddm_opt;
  FILE *log = (ho->hddm_log == REF_NEED) ? ho->fp : NULL;
  fwh_ctx *ctx = NULL;
  fwh_stats st;
  memset(&st, 0, sizeof(st));
  double *x = NULL, *nu = NULL, *B = NULL, *E = NULL, *J = NULL, *res = NULL;
  if ((rc = fwh_create(&ctx, device_from_env())) || (rc = fwh_load(ctx, &d))) goto done;
  {
    fwh_info info;
    if ((rc = fwh_get_info(ctx, &info))) goto done;
    /* hddm_solver.h:86-90 (Fprintf prefixes "#", advfullwave_lib.h:56) */
    if (log) fprintf(log, "# DOF on interface %d\n", info.n_iface);
    /* the batched form: explicit S_s on the matrix cores, iterations in
     * hipGraphs; the stop test and status machine are the reference's    */
    fwh_opts o;
    memset(&o, 0, sizeof(o));
    o.hddm_eps = ho->hddm_eps;
    o.hddm_max_loop = ho->hddm_max_loop;
    o.plain = 0;
    o.graph_iters = 16;
    o.schur = 1;
    if ((rc = fwh_solve(ctx, &o, &st))) goto done;
    if (ho->hddm_max_loop == 0) goto done; /* not executed: no results (hddm_solver.h:103-105) */
    if (log) {
      /* the residual history in the reference's format (hddm_solver_s.h:
       * 205, 321): loop 0 with the initial norm, then one line per loop   */
      res = (double *)zalloc(st.loops, sizeof(double));
      if ((rc = fwh_get_resid(ctx, res, st.loops))) goto done;
      fprintf(log, "%d %le  # Norm %le\n", 0, 1.0, st.r0);
      for (int k = 1; k < st.loops; k++) fprintf(log, "%d %le\n", k, res[k]);
      fflush(log);
    }
    /* get_result_sp (fullwave.c:500-514): dm->x_c of every subdomain;
     * post_sp (fullwave.c:607-635): dm->rvalue.el_{B,E,Je}_c             */
    x = (double *)zalloc(2 * (size_t)info.n_dof, sizeof(double));
    nu = (double *)zalloc(info.n_elem, sizeof(double));
    B = (double *)zalloc(6 * (size_t)info.n_elem, sizeof(double));
    E = (double *)zalloc(6 * (size_t)info.n_elem, sizeof(double));
    J = (double *)zalloc(6 * (size_t)info.n_elem, sizeof(double));
    if (!x || !nu || !B || !E || !J) {
      rc = ref_fail(FWH_E_ARG, "out of host memory");
      goto done;
    }
    {
      long el = 0;
      for (int p = 0; p < d.npart; p++)
        for (int k = 0; k < pm[p].ndomain; k++)
          for (int e = 0; e < pm[p].dm[k].nel; e++, el++)
            nu[el] = has_flag(&pm[p].mtrl, pm[p].dm[k].flag[e]) ? REF_NU : 0.0;
    }
    if ((rc = fwh_get_x(ctx, x)) || (rc = fwh_post(ctx, nu, B, E, J))) goto done;
    {
      size_t xo = 0, eo = 0;
      for (int p = 0; p < d.npart; p++)
        for (int k = 0; k < pm[p].ndomain; k++) {
          DomMesh *dm = &pm[p].dm[k];
          const size_t nm = (size_t)dm->nmp, ne = (size_t)dm->nel;
          Complex *xc = (Complex *)zalloc(nm, sizeof(Complex));
          Complex *bc = (Complex *)zalloc(3 * ne, sizeof(Complex));
          Complex *ec = (Complex *)zalloc(3 * ne, sizeof(Complex));
          Complex *jc = (Complex *)zalloc(3 * ne, sizeof(Complex));
          if (!xc || !bc || !ec || !jc) {
            free(xc); free(bc); free(ec); free(jc);
            rc = ref_fail(FWH_E_ARG, "out of host memory");
            goto done;
          }
          memcpy(xc, x + 2 * xo, nm * sizeof(Complex));
          memcpy(bc, B + 6 * eo, 3 * ne * sizeof(Complex));
          memcpy(ec, E + 6 * eo, 3 * ne * sizeof(Complex));
          memcpy(jc, J + 6 * eo, 3 * ne * sizeof(Complex));
          dm->x_c = xc;
          dm->rvalue.el_B_c = bc;
          dm->rvalue.el_E_c = ec;
          dm->rvalue.el_Je_c = jc;
          xo += nm;
          eo += ne;
        }
    }
  }
done:
  if (stats) *stats = st;
  free(x); free(nu); free(B); free(E); free(J); free(res);
  if (ctx) fwh_destroy(ctx);
  fwh_ref_desc_free(&d);
  return rc;
}

void hddm_solver_CS_hip(PartMesh *pm, OPTIONS options) {
  const int rc = hddm_solver_CS_hip_status(pm, options, NULL);
  if (rc) {
    /* emessage(code, abort, str) -> COM_Exit (lib/std/errormessage.h:13) */
    fprintf(stdout, "ERROR: hddm_solver_CS_hip (%d): %s\n", rc, fwh_last_error());
    fflush(stdout);
    exit(1);
  }
}

/* ---- the skyline calls on a process-wide context ----------------------- */
static fwh_ctx *g_sky_ctx = NULL;

static fwh_ctx *sky_ctx(void) {
  if (!g_sky_ctx && fwh_create(&g_sky_ctx, device_from_env()) != 0) {
    fprintf(stdout, "ERROR: libfwhip skyline context: %s\n", fwh_last_error());
    fflush(stdout);
    exit(1);
  }
  return g_sky_ctx;
}

void advlas_skyline_ldl_decomposite(DMatrix dmat, double _Complex *work) {
  const int rc = fwh_sky_factor(sky_ctx(), dmat.ndim, dmat.idx1, dmat.idx2, (double *)dmat.mat_dc, (double *)work);
  if (rc) {
    fprintf(stdout, "ERROR: advlas_skyline_ldl_decomposite (%d): %s\n", rc, fwh_last_error());
    fflush(stdout);
    exit(1);
  }
}

void advlas_skyline_ldl_solve(DMatrix dmat, double _Complex *solution) {
  const int rc = fwh_sky_solve(sky_ctx(), dmat.ndim, dmat.idx1, dmat.idx2, (const double *)dmat.mat_dc,
                               (double *)solution);
  if (rc) {
    fprintf(stdout, "ERROR: advlas_skyline_ldl_solve (%d): %s\n", rc, fwh_last_error());
    fflush(stdout);
    exit(1);
  }
}

/* ---- the subdomain COCG of mymt_aij_solver on the GPU ------------------ */
/* mymt_CS_AIJ_Solver_COCG (mymt_aij_cs.h:64, aij_solver_cg.h:9-44): the
 * preconditioner switch, then COCG on aij_mat->x_c from aij_mat->b_c.  ICC
 * and none run on the GPU (fwh_aij_cocg_pc); the IC(0) factor is left in
 * aij_mat->pc_c for the _nmpcm form, as the reference's make step leaves
 * it; MRIC0 and the others return SOLVER_STATUS_UNKNOWN_PC, as the
 * reference's switch does for a PC it does not know.  The monitor log is
 * not written. */
/* the reference's counted allocator (lib/std/mem.c:86) when the library is
 * linked into the reference, else malloc: the IC(0) factor left in
 * aij_mat->pc_c is released by mymt_CS_AIJ_Free (aij_matrix_func.h:253-255)
 * with the reference's Free, i.e. free() */
extern void *__MemoryAllocate__(size_t mem, char *name, char *fname, int line) __attribute__((weak));

static Complex *pc_alloc(size_t count) {
  const size_t bytes = sizeof(Complex) * count;
  return (Complex *)(__MemoryAllocate__ ? __MemoryAllocate__(bytes, NULL, (char *)__FILE__, __LINE__)
                                        : malloc(bytes));
}

static int aij_cocg(AIJ_MAT *aij_mat, SolvOpt solvopt, int make, const char *who) {
  const int nf = aij_mat->nf;
  const int pc_size = aij_mat->ai[nf] + nf;
  double *pcm = NULL;
  if (solvopt.solver_pc == REF_SOLVER_PC_ICC) {
    if (make) {
      /* PC_ICC make (aij_pc_func.h:52-58): a factor of ai[nf] + nf
       * entries in aij_mat->pc_c (a buffer of that size is reused) */
      if (!aij_mat->pc_c || aij_mat->pc_size != pc_size) {
        aij_mat->pc_c = pc_alloc((size_t)pc_size);
        if (!aij_mat->pc_c) {
          fprintf(stdout, "ERROR: %s: out of host memory for the preconditioner\n", who);
          fflush(stdout);
          exit(1);
        }
      }
      aij_mat->pc_size = pc_size;
    } else if (!aij_mat->pc_c || aij_mat->pc_size != pc_size) {
      fprintf(stdout, "ERROR: %s: no preconditioner matrix of %d entries in aij_mat\n", who, pc_size);
      fflush(stdout);
      exit(1);
    }
    pcm = (double *)aij_mat->pc_c;
  }
  int status = 0;
  const int rc = fwh_aij_cocg_pc(sky_ctx(), nf, aij_mat->ai, aij_mat->aij, (const double *)aij_mat->vij_c,
                                 (const double *)aij_mat->b_c, (double *)aij_mat->x_c, solvopt.solver_pc,
                                 solvopt.solver_eps, solvopt.solver_div, solvopt.solver_ic_shift, pcm, make, &status);
  if (rc) {
    fprintf(stdout, "ERROR: %s (%d): %s\n", who, rc, fwh_last_error());
    fflush(stdout);
    exit(1);
  }
  return status;
}

int mymt_CS_AIJ_Solver_COCG(AIJ_MAT *aij_mat, SolvOpt solvopt, FILE *monitor) {
  (void)monitor;
  if (solvopt.solver_pc != REF_SOLVER_PC_NONE && solvopt.solver_pc != REF_SOLVER_PC_ICC)
    return REF_SOLVER_STATUS_UNKNOWN_PC;
  return aij_cocg(aij_mat, solvopt, 1, "mymt_CS_AIJ_Solver_COCG");
}

/* mymt_CS_AIJ_Solver_COCG_nmpcm (mymt_aij_cs.h:67, aij_solver_cg.h:68-92):
 * the same COCG with the preconditioner the make step left in aij_mat
 * (pc_c, pc_size) for SOLVER_PC_ICC, none for SOLVER_PC_NONE,
 * SOLVER_STATUS_UNKNOWN_PC for the others. */
int mymt_CS_AIJ_Solver_COCG_nmpcm(AIJ_MAT *aij_mat, SolvOpt solvopt, FILE *monitor) {
  (void)monitor;
  if (solvopt.solver_pc != REF_SOLVER_PC_NONE && solvopt.solver_pc != REF_SOLVER_PC_ICC)
    return REF_SOLVER_STATUS_UNKNOWN_PC;
  return aij_cocg(aij_mat, solvopt, 0, "mymt_CS_AIJ_Solver_COCG_nmpcm");
}
