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
 * set_hddm_func_fullwave_hip fills the reference's hook table
 * (lib/std/types.h:629-730) so that the reference's own driver loop runs the
 * same batched GPU solve: the per-subdomain hooks queue, the iteration
 * hooks run the work for all subdomains.
 *
 * advlas_skyline_ldl_decomposite / _solve keep the reference's prototypes
 * (lib/hddm/advlas_skyline.h:11-12) and run on the GPU of a process-wide
 * context, as do mymt_CS_AIJ_Solver_COCG / _nmpcm (mymt_aij_cs.h:64-68).
 */
#define _GNU_SOURCE /* RTLD_DEFAULT */
#include <complex.h>
#include <dlfcn.h>
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

/* ---- the hook table: set_hddm_func_fullwave_hip ------------------------ */
#define REF_HDDM_SW_ITERATION 1 /* __HDDM_SW_ITERATION__, lib/hddm/libhddm.h:38 */
#define REF_HDDM_SW_LAST 2      /* __HDDM_SW_LAST__, lib/hddm/libhddm.h:39 */
#define REF_HDDM_PC_DIAG 1      /* HDDM_PC_DIAG, lib/std/defs.h */
#define HOOK_GRAPH 16           /* interface iterations per hipGraph */

/* the reference's own, when the library is linked into it: the solver
 * state (lib/hddm/make_PartMesh.c:16), the full-wave hook table
 * (main/fullwave/fullwave.c:23-146) and the counted Calloc (lib/std/mem.c) */
extern HDDM_SW hddm_sw __attribute__((weak));
extern void set_hddm_func_fullwave(HDDM_FUNC *hddm_func, OPTIONS options) __attribute__((weak));
extern void *__ClearAllocate__(size_t mem, char *name, char *fname, int line) __attribute__((weak));

/* host arrays the reference frees with its Free (free()) */
static void *ref_calloc(size_t n, size_t size) {
  if (n == 0) n = 1;
  if (__ClearAllocate__) return __ClearAllocate__(n * size, NULL, (char *)__FILE__, __LINE__);
  return calloc(n, size);
}

/* one model per process, from general_func_1st to general_func_5th */
static struct {
  fwh_ctx *ctx;
  fwh_mesh_desc d;
  int have_desc;
  PartMesh *pm;   /* the parts general_func_1st was given */
  int npart;
  int n_iface;
  long *xoff;     /* per part: its first DOF in the flat solution (npart + 1) */
  long *eoff;     /* per part: its first element (npart + 1) */
  fwh_opts o;
  int calls;      /* iteration_s calls since iteration_s_1st */
  int run;        /* interface iterations launched on the GPU */
  int done;       /* the GPU's status is LAST ... */
  int k_stop;     /* ... since the iteration of this loop */
  double *hist;   /* residual history read back so far (hist_n entries) */
  int hist_n;
  double *x, *B, *E, *J; /* the results, read back once */
} H;

static void hooks_release(void) {
  if (H.ctx) fwh_destroy(H.ctx);
  if (H.have_desc) fwh_ref_desc_free(&H.d);
  free(H.xoff); free(H.eoff); free(H.hist);
  free(H.x); free(H.B); free(H.E); free(H.J);
  memset(&H, 0, sizeof(H));
}

static void hook_fail(const char *who, int rc) {
  /* emessage(code, abort, str) -> COM_Exit (lib/std/errormessage.h:13) */
  fprintf(stdout, "ERROR: %s (%d): %s\n", who, rc, fwh_last_error());
  fflush(stdout);
  exit(1);
}

static void hook_check(const char *who, int rc) {
  if (rc) hook_fail(who, rc);
}

/* the reference's hddm_sw: the library's own reference to it when the
 * process defines it at load, else a symbol of the process (a driver
 * loaded later with global symbols) */
static HDDM_SW *ref_sw(const char *who) {
  HDDM_SW *sw = &hddm_sw ? &hddm_sw : (HDDM_SW *)dlsym(RTLD_DEFAULT, "hddm_sw");
  if (!sw) {
    fprintf(stdout, "ERROR: %s: the reference's hddm_sw is not a symbol of this process\n", who);
    fflush(stdout);
    exit(1);
  }
  return sw;
}

static int hook_part(const PartMesh *pm, const char *who) {
  const long p = H.pm ? (long)(pm - H.pm) : -1;
  if (!H.ctx || p < 0 || p >= H.npart) {
    fprintf(stdout, "ERROR: %s: not a part of the model general_func_1st loaded\n", who);
    fflush(stdout);
    exit(1);
  }
  return (int)p;
}

/* general_func_1st (hddm_solver.h:77): the whole model onto the GPU, as
 * hddm_solver_CS_hip loads it */
static void hk_general_1st(PartMesh *pm, HDDM_VEC *v, OPTIONS options, HDDM_Opt ho, SolvOpt so) {
  (void)v; (void)options; (void)ho; (void)so;
  hooks_release();
  hook_check("general_func_1st", fwh_ref_desc_from_partmesh(pm, &H.d));
  H.have_desc = 1;
  hook_check("general_func_1st", fwh_create(&H.ctx, device_from_env()));
  hook_check("general_func_1st", fwh_load(H.ctx, &H.d));
  fwh_info info;
  hook_check("general_func_1st", fwh_get_info(H.ctx, &info));
  H.pm = pm;
  H.npart = H.d.npart;
  H.n_iface = info.n_iface;
  H.xoff = (long *)zalloc(H.npart + 1, sizeof(long));
  H.eoff = (long *)zalloc(H.npart + 1, sizeof(long));
  if (!H.xoff || !H.eoff) hook_fail("general_func_1st", ref_fail(FWH_E_ARG, "out of host memory"));
  for (int p = 0; p < H.npart; p++) {
    long nx = 0, ne = 0;
    for (int k = 0; k < pm[p].ndomain; k++) {
      nx += pm[p].dm[k].nmp;
      ne += pm[p].dm[k].nel;
    }
    H.xoff[p + 1] = H.xoff[p] + nx;
    H.eoff[p + 1] = H.eoff[p] + ne;
  }
}

/* general_func_5th (hddm_solver.h:157): the GPU released once the results
 * are in the DomMesh structs */
static void hk_general_5th(PartMesh *pm, HDDM_VEC *v, OPTIONS options, HDDM_Opt ho, SolvOpt so) {
  (void)pm; (void)v; (void)options; (void)ho; (void)so;
  hooks_release();
}

static void hk_general_none(PartMesh *pm, HDDM_VEC *v, OPTIONS options, HDDM_Opt ho, SolvOpt so) {
  (void)pm; (void)v; (void)options; (void)ho; (void)so;
}

/* make_HDDM_VEC (fullwave.c:168-174, hddm_solver.h:292-313): the part's
 * host vectors, sized as the reference sizes them (the iterates themselves
 * live on the GPU) */
static void hk_make_vec(HDDM_VEC *v, PartMesh pm, HDDM_Opt ho) {
  v->infree = pm.infree;
  v->outfree = pm.outfree;
  v->tfree = pm.infree + pm.outfree;
  v->total_free = pm.total_free > 0 ? pm.total_free : H.n_iface;
  v->p_c = (Complex *)ref_calloc(v->tfree, sizeof(Complex));
  v->q_c = (Complex *)ref_calloc(v->tfree, sizeof(Complex));
  v->r_c = (Complex *)ref_calloc(v->infree, sizeof(Complex));
  v->z_c = (Complex *)ref_calloc(v->infree, sizeof(Complex));
  v->u_c = (Complex *)ref_calloc(v->infree, sizeof(Complex));
  if (ho.hddm_pc == REF_HDDM_PC_DIAG) {
    v->pc_size = v->tfree;
    v->pc_c = (Complex *)ref_calloc(v->tfree, sizeof(Complex));
  }
}

/* the per-subdomain hooks of the loops: their work is queued for the
 * iteration hooks, which run it for every subdomain at once */
static void hk_dom_none(DomMesh *dm) { (void)dm; }
static void hk_make_tmp_aij(DomMesh *dm, MTRL mtrl) { (void)dm; (void)mtrl; }
static void hk_renewal(DomMesh *dm, HDDM_VEC v) { (void)dm; (void)v; }
static void hk_solve_dom_1st(DomMesh *dm, HDDM_Opt ho, SolvOpt so, fwh_ref_dom_solver_fn ds, HDDM_DOM_VEC *dv,
                             fwh_ref_make_pc_fn mk, FILE *fp) {
  (void)dm; (void)ho; (void)so; (void)ds; (void)dv; (void)mk; (void)fp;
}
static void hk_solve_dom(DomMesh *dm, SolvOpt so, fwh_ref_dom_solver_fn ds, HDDM_DOM_VEC *dv, FILE *fp) {
  (void)dm; (void)so; (void)ds; (void)dv; (void)fp;
}
static void hk_com_res(HDDM_VEC *v, HDDM_DOM_VEC dv, DomMesh dm, HDDM_Opt ho) {
  (void)v; (void)dv; (void)dm; (void)ho;
}
static int hk_make_pc(TMP_AIJ t, HDDM_DOM_VEC *dv) { (void)t; (void)dv; return 0; }
static int hk_set_pc(HDDM_VEC *v) { (void)v; return 0; }
static void hk_gather_pc(OPinfo **oop, HDDM_VEC *v) { (void)oop; (void)v; }

/* iteration_s_1st (hddm_solver_s.h:158-256): the first step of every
 * subdomain and the interface residual r0; status ITERATION, or LAST when
 * ||r0|| < NEARLY_ZERO */
static void hk_iteration_1st(PartMesh *pm, HDDM_VEC *v, HDDM_Opt ho, fwh_ref_set_pc_fn setPC,
                             void (*gatherPC_s)(OPinfo **oop, HDDM_VEC *hddm_vec)) {
  (void)v; (void)setPC; (void)gatherPC_s;
  (void)hook_part(pm, "iteration_s_1st");
  HDDM_SW *sw = ref_sw("iteration_s_1st");
  memset(&H.o, 0, sizeof(H.o));
  H.o.hddm_eps = ho.hddm_eps;
  H.o.hddm_max_loop = ho.hddm_max_loop;
  H.o.plain = 0;
  H.o.graph_iters = HOOK_GRAPH;
  H.o.schur = 1;
  int done = 0;
  fwh_stats st;
  hook_check("iteration_s_1st", fwh_first(H.ctx, &H.o));
  hook_check("iteration_s_1st", fwh_iterate(H.ctx, &H.o, 0, &done));
  hook_check("iteration_s_1st", fwh_get_stats(H.ctx, &st));
  if (ho.hddm_log == REF_NEED && ho.fp) fprintf(ho.fp, "%d %le  # Norm %le\n", 0, 1.0, st.r0);
  H.calls = 0;
  H.run = 0;
  H.done = done;
  H.k_stop = 0;
  sw->status = done ? REF_HDDM_SW_LAST : REF_HDDM_SW_ITERATION;
}

/* iteration_s (hddm_solver_s.h:260-400), the k-th call of the loop: the
 * GPU runs ahead in graphs of HOOK_GRAPH iterations (its status machine
 * makes the iterations after the stop no-ops); the call logs the k-th
 * residual and sets LAST at the loop where the GPU stopped (convergence,
 * or hddm_max_loop, where the reference's loop sets it too) */
static void hk_iteration(PartMesh *pm, HDDM_VEC *v, HDDM_Opt ho, fwh_ref_set_pc_fn setPC) {
  (void)v; (void)setPC;
  (void)hook_part(pm, "iteration_s");
  HDDM_SW *sw = ref_sw("iteration_s");
  const int k = ++H.calls;
  if (!H.done && k > H.run) {
    int done = 0;
    fwh_stats st;
    hook_check("iteration_s", fwh_iterate(H.ctx, &H.o, HOOK_GRAPH, &done));
    H.run += HOOK_GRAPH;
    hook_check("iteration_s", fwh_get_stats(H.ctx, &st));
    if (st.loops > H.hist_n) {
      double *h = (double *)realloc(H.hist, sizeof(double) * (size_t)st.loops);
      if (!h) hook_fail("iteration_s", ref_fail(FWH_E_ARG, "out of host memory"));
      H.hist = h;
      hook_check("iteration_s", fwh_get_resid(H.ctx, H.hist, st.loops));
      H.hist_n = st.loops;
    }
    if (done) {
      H.done = 1;
      H.k_stop = st.loops - 1;
    }
  }
  if (ho.hddm_log == REF_NEED && ho.fp && k < H.hist_n) fprintf(ho.fp, "%d %le\n", k, H.hist[k]);
  if (H.done && k >= H.k_stop) sw->status = REF_HDDM_SW_LAST;
}

/* get_result_sp (fullwave.c:500-514): after the last step, dm->x_c of the
 * part's subdomains */
static void hk_get_result(PartMesh *pm) {
  const int p = hook_part(pm, "get_result_sp");
  if (!H.x) {
    fwh_info info;
    hook_check("get_result_sp", fwh_last(H.ctx));
    hook_check("get_result_sp", fwh_get_info(H.ctx, &info));
    H.x = (double *)zalloc(2 * (size_t)info.n_dof, sizeof(double));
    if (!H.x) hook_fail("get_result_sp", ref_fail(FWH_E_ARG, "out of host memory"));
    hook_check("get_result_sp", fwh_get_x(H.ctx, H.x));
  }
  long xo = H.xoff[p];
  for (int k = 0; k < pm->ndomain; k++) {
    DomMesh *dm = &pm->dm[k];
    dm->x_c = (Complex *)ref_calloc((size_t)dm->nmp, sizeof(Complex));
    if (!dm->x_c) hook_fail("get_result_sp", ref_fail(FWH_E_ARG, "out of host memory"));
    memcpy(dm->x_c, H.x + 2 * xo, (size_t)dm->nmp * sizeof(Complex));
    xo += dm->nmp;
  }
}

/* post_sp (fullwave.c:607-635): dm->rvalue.el_B_c / el_E_c / el_Je_c of
 * the part's elements */
static void hk_post(PartMesh *pm) {
  const int p = hook_part(pm, "post_sp");
  if (!H.x) hook_fail("post_sp", ref_fail(FWH_E_STATE, "post_sp before get_result_sp"));
  if (!H.B) {
    fwh_info info;
    hook_check("post_sp", fwh_get_info(H.ctx, &info));
    const size_t ne = (size_t)info.n_elem;
    double *nu = (double *)zalloc(ne, sizeof(double));
    H.B = (double *)zalloc(6 * ne, sizeof(double));
    H.E = (double *)zalloc(6 * ne, sizeof(double));
    H.J = (double *)zalloc(6 * ne, sizeof(double));
    if (!nu || !H.B || !H.E || !H.J) {
      free(nu);
      hook_fail("post_sp", ref_fail(FWH_E_ARG, "out of host memory"));
    }
    long el = 0;
    for (int q = 0; q < H.npart; q++)
      for (int k = 0; k < H.pm[q].ndomain; k++)
        for (int e = 0; e < H.pm[q].dm[k].nel; e++, el++)
          nu[el] = has_flag(&H.pm[q].mtrl, H.pm[q].dm[k].flag[e]) ? REF_NU : 0.0;
    const int rc = fwh_post(H.ctx, nu, H.B, H.E, H.J);
    free(nu);
    hook_check("post_sp", rc);
  }
  long eo = H.eoff[p];
  for (int k = 0; k < pm->ndomain; k++) {
    DomMesh *dm = &pm->dm[k];
    const size_t ne = (size_t)dm->nel;
    dm->rvalue.el_B_c = (Complex *)ref_calloc(3 * ne, sizeof(Complex));
    dm->rvalue.el_E_c = (Complex *)ref_calloc(3 * ne, sizeof(Complex));
    dm->rvalue.el_Je_c = (Complex *)ref_calloc(3 * ne, sizeof(Complex));
    if (!dm->rvalue.el_B_c || !dm->rvalue.el_E_c || !dm->rvalue.el_Je_c)
      hook_fail("post_sp", ref_fail(FWH_E_ARG, "out of host memory"));
    memcpy(dm->rvalue.el_B_c, H.B + 6 * eo, 3 * ne * sizeof(Complex));
    memcpy(dm->rvalue.el_E_c, H.E + 6 * eo, 3 * ne * sizeof(Complex));
    memcpy(dm->rvalue.el_Je_c, H.J + 6 * eo, 3 * ne * sizeof(Complex));
    eo += (long)ne;
  }
}

/* alone (not linked into the reference): options copied, no result files */
static void hk_set_hddm_opt(HDDM_Opt *ho, OPTIONS options) { *ho = options.hddm_opt; }
static void hk_set_solvopt(SolvOpt *so, OPTIONS options) { *so = options.solvopt; }
static void hk_out_result(int t_loop, PartMesh pm, OPTIONS options) { (void)t_loop; (void)pm; (void)options; }

void set_hddm_func_fullwave_hip(HDDM_FUNC *f, OPTIONS options) {
  if (set_hddm_func_fullwave) {
    set_hddm_func_fullwave(f, options);
  } else {
    memset(f, 0, sizeof(*f));
    f->set_HDDM_Opt = hk_set_hddm_opt;
    f->set_SolvOpt = hk_set_solvopt;
    f->free_TMP_AIJ = hk_dom_none;
    f->out_result = hk_out_result;
    f->general_func_2nd = hk_general_none;
    f->general_func_3rd = hk_general_none;
    f->general_func_4th = hk_general_none;
  }
  f->general_func_1st = hk_general_1st;
  f->general_func_5th = hk_general_5th;
  f->make_HDDM_VEC = hk_make_vec;
  f->init_bc = hk_dom_none;
  f->make_TMP_AIJ = hk_make_tmp_aij;
  f->renewal_bc_sp = hk_renewal;
  f->dom_solver = mymt_CS_AIJ_Solver_COCG;
  f->dom_solver_nmpcm = mymt_CS_AIJ_Solver_COCG_nmpcm;
  f->makePC = hk_make_pc;
  f->setPC = hk_set_pc;
  f->gatherPC_s = hk_gather_pc;
  f->solve_dom_1st = hk_solve_dom_1st;
  f->solve_dom = hk_solve_dom;
  f->solve_dom_last = hk_solve_dom;
  f->COM_res_sp_1st = hk_com_res;
  f->COM_res_sp = hk_com_res;
  f->iteration_s_1st = hk_iteration_1st;
  f->iteration_s = hk_iteration;
  f->get_result_sp = hk_get_result;
  f->post_sp = hk_post;
}
