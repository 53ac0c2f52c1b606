/*
 * fw_oracle.c - CPU restatement of the ADVENTURE_FullWave HDDM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark.
 *
 * Written from the reference's algorithms; each block names the reference
 * routine (file:line) whose semantics it follows.  Arithmetic follows the
 * reference: struct complex with the multiply / naive reciprocal of
 * lib/std/complex_adv.c:39-137 in the sparse and HDDM code, and C99
 * double _Complex in the skyline factor/solve (lib/hddm/advlas_skyline.c).
 */
#include "fw_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NEARLY_ZERO 1.0e-60 /* lib/std/defs.h:105 */

/* ------------------------------------------------------------------------ */
/* complex arithmetic, lib/std/complex_adv.c                                */
/* ------------------------------------------------------------------------ */
static const fwo_c C0 = {0.0, 0.0};
static const fwo_c C1 = {1.0, 0.0};

static inline fwo_c cadd(fwo_c a, fwo_c b) { fwo_c r = {a.re + b.re, a.im + b.im}; return r; }
static inline fwo_c csub(fwo_c a, fwo_c b) { fwo_c r = {a.re - b.re, a.im - b.im}; return r; }
/* ComplexMulti, complex_adv.c:103-114 */
static inline fwo_c cmul(fwo_c a, fwo_c b) {
  fwo_c r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.im * b.re + a.re * b.im;
  return r;
}
static inline fwo_c cscale(fwo_c a, double s) { fwo_c r = {a.re * s, a.im * s}; return r; }
/* ComplexReciprocal, complex_adv.c:39-51 */
static inline fwo_c crecip(fwo_c a) {
  double d = 1.0 / (a.re * a.re + a.im * a.im);
  fwo_c r = {a.re * d, -(a.im * d)};
  return r;
}
/* ComplexDivide = a * recip(b), complex_adv.c:133-137 */
static inline fwo_c cdiv(fwo_c a, fwo_c b) { return cmul(a, crecip(b)); }
/* ComplexAbs, complex_adv.c:56-60 */
static inline double cabs_(fwo_c a) { return sqrt(a.re * a.re + a.im * a.im); }
static inline int ciszero(fwo_c a) { return a.re == 0.0 && a.im == 0.0; }

/* ComplexBilinear (aTb, four partial sums), complex_adv.c:185-207 */
static fwo_c cbilinear(int n, const fwo_c *a, const fwo_c *b) {
  double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
  for (int i = 0; i < n; i++) {
    d0 += a[i].re * b[i].re;
    d1 += a[i].im * b[i].im;
    d2 += a[i].re * b[i].im;
    d3 += a[i].im * b[i].re;
  }
  fwo_c r = {d0 - d1, d2 + d3};
  return r;
}
/* ComplexInner (aHb), complex_adv.c:160-181 */
static fwo_c cinner(int n, const fwo_c *a, const fwo_c *b) {
  double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
  for (int i = 0; i < n; i++) {
    d0 += a[i].re * b[i].re;
    d1 += a[i].im * b[i].im;
    d2 += a[i].re * b[i].im;
    d3 += a[i].im * b[i].re;
  }
  fwo_c r = {d0 + d1, d2 - d3};
  return r;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* element kernel, lib/fem/make_elem_em.c:14-400 with lib/fem/tetrahedral.c */
/* ------------------------------------------------------------------------ */

/* Local edge k joins vertices EDGE[k][0] -> EDGE[k][1]; EDGE[k][2..3] is the
 * opposite edge, oriented so that grad(l_a) x grad(l_b) = (v_d - v_c) / D.
 * Same table as the reference's tet_n (lib/fem/tetrahedral.h:19-24): the
 * mesh format fixes the edge numbering.                                    */
static const int EDGE[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                               {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};

static const double NU = 7.957747e+05; /* fixed 1/mu0, make_elem_em.c:37 */

/* Geometry shared by the element routines: G[i] = D grad(lambda_i) (the
 * reference's simple_b/c/d, tetrahedral.c:122-184), D = 6V
 * (tet_Volume6, :44-61), centroid g (tet_Center, :80-86), and per edge k
 * the signed length ls_k = l_k si_k (tet_SideLength + tetNedelec_Direction,
 * :27-40,225-263), e_k = v_d - v_c (tetNedelec_mk, :267-281) and
 * w_k = v_c x v_d (tetNedelec_km_mk, :305-319).  Returns <0 for an
 * inverted or degenerate element.                                         */
static int tet_frame(const double *v, double G[4][3], double *Dout, double g[3], double ls[6],
                     double e[6][3], double w[6][3]) {
  double r1[3], r2[3], r3[3];
  for (int a = 0; a < 3; a++) {
    r1[a] = v[3 + a] - v[a];
    r2[a] = v[6 + a] - v[a];
    r3[a] = v[9 + a] - v[a];
  }
  /* columns of adj(J), J rows r1, r2, r3 */
  G[1][0] = r2[1] * r3[2] - r2[2] * r3[1];
  G[1][1] = r2[2] * r3[0] - r2[0] * r3[2];
  G[1][2] = r2[0] * r3[1] - r2[1] * r3[0];
  G[2][0] = r3[1] * r1[2] - r3[2] * r1[1];
  G[2][1] = r3[2] * r1[0] - r3[0] * r1[2];
  G[2][2] = r3[0] * r1[1] - r3[1] * r1[0];
  G[3][0] = r1[1] * r2[2] - r1[2] * r2[1];
  G[3][1] = r1[2] * r2[0] - r1[0] * r2[2];
  G[3][2] = r1[0] * r2[1] - r1[1] * r2[0];
  for (int a = 0; a < 3; a++) G[0][a] = -(G[1][a] + G[2][a] + G[3][a]);
  double D = r1[0] * G[1][0] + r1[1] * G[1][1] + r1[2] * G[1][2];
  if (!(D > 0.0)) return -1;
  *Dout = D;
  for (int a = 0; a < 3; a++) g[a] = (v[a] + v[3 + a] + v[6 + a] + v[9 + a]) * 0.25;
  for (int k = 0; k < 6; k++) {
    const double *va = v + 3 * EDGE[k][0], *vb = v + 3 * EDGE[k][1];
    const double *vc = v + 3 * EDGE[k][2], *vd = v + 3 * EDGE[k][3];
    double dl[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
    double l = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    int si = 0;
    for (int a = 0; a < 3 && si == 0; a++) si = dl[a] > 0.0 ? 1 : (dl[a] < 0.0 ? -1 : 0);
    if (si == 0) return -2;
    ls[k] = l * (double)si;
    for (int a = 0; a < 3; a++) e[k][a] = vd[a] - vc[a];
    w[k][0] = vc[1] * vd[2] - vc[2] * vd[1];
    w[k][1] = vc[2] * vd[0] - vc[0] * vd[2];
    w[k][2] = vc[0] * vd[1] - vc[1] * vd[0];
  }
  return 0;
}

int fwo_element(const double *v, double eps, double sigma, double omega,
                const double *pml, int src, const double *joe, double *ae72,
                double *be12) {
  double G[4][3], D, g[3], ls[6], e[6][3], w[6][3];
  int rc = tet_frame(v, G, &D, g, ls, e, w);
  if (rc != 0) return rc;
  /* N_k(r) = (ls_k / D) (w_k + e_k x r); F_k is the bracket at the
   * centroid (tetNedelec_Fxyz, tetrahedral.c:323-340, without ls)        */
  double F[6][3];
  for (int k = 0; k < 6; k++) {
    F[k][0] = w[k][0] + e[k][1] * g[2] - e[k][2] * g[1];
    F[k][1] = w[k][1] + e[k][2] * g[0] - e[k][0] * g[2];
    F[k][2] = w[k][2] + e[k][0] * g[1] - e[k][1] * g[0];
  }
  /* second moments about the centroid: C = (1/V) int rho rho^T dV
   * = (1/20) sum_i (v_i - g)(v_i - g)^T                                    */
  double C[3][3] = {{0}};
  for (int i = 0; i < 4; i++) {
    double p[3] = {v[3 * i] - g[0], v[3 * i + 1] - g[1], v[3 * i + 2] - g[2]};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) C[a][b] += p[a] * p[b];
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) C[a][b] *= 0.05;

  /* per-component curl-curl (tetNedelec_rotrot, tetrahedral.c:344-358)
   * RR^a_ij = 2/(3D) ls_i ls_j e_i,a e_j,a, and per-component mass
   * (tetNedelec_sideside, tetrahedral.c:362-422)
   * M^a_ij = ls_i ls_j /(6D) (F_i,a F_j,a + E[(e_i x rho)_a (e_j x rho)_a]) */
  const double drr = 2.0 / (3.0 * D), dss = 1.0 / (6.0 * D);
  double RR[3][6][6], MM[3][6][6];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3; /* (e x rho)_a = e_b rho_c - e_c rho_b */
        double q = e[i][b] * e[j][b] * C[c][c] - e[i][b] * e[j][c] * C[c][b] -
                   e[i][c] * e[j][b] * C[b][c] + e[i][c] * e[j][c] * C[b][b];
        RR[a][i][j] = drr * ls[i] * ls[j] * e[i][a] * e[j][a];
        MM[a][i][j] = dss * ls[i] * ls[j] * (F[i][a] * F[j][a] + q);
      }
    }

  fwo_c *ae = (fwo_c *)ae72;
  fwo_c *be = (fwo_c *)be12;
  const double nu_omega = NU / omega; /* make_elem_em.c:227 */
  const double dd = -(eps * omega);   /* make_elem_em.c:239 */
  if (pml == NULL) {
    /* ae = (nu/w) RR - (eps w) SS + i sigma SS, accumulated as the
     * reference does: re += nu_omega*rr; re += dd*ss; im += sigma*ss      */
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) {
        double rr = RR[0][i][j] + RR[1][i][j] + RR[2][i][j];
        double ss = MM[0][i][j] + MM[1][i][j] + MM[2][i][j];
        fwo_c t = C0;
        t.re += nu_omega * rr;
        t.re += dd * ss;
        t.im += sigma * ss;
        ae[i * 6 + j] = t;
      }
  } else {
    /* uniaxial stretching: mu^-1 -> Lambda^-1 mu^-1, eps_c -> Lambda eps_c
     * with Lambda = diag(sy sz/sx, sx sz/sy, sx sy/sz).  No reference
     * counterpart; stretch 1 reduces to the branch above.                 */
    fwo_c s[3] = {{pml[0], pml[1]}, {pml[2], pml[3]}, {pml[4], pml[5]}};
    fwo_c lam[3], ilam[3];
    for (int a = 0; a < 3; a++) {
      int b = (a + 1) % 3, c = (a + 2) % 3;
      lam[a] = cdiv(cmul(s[b], s[c]), s[a]);
      ilam[a] = crecip(lam[a]);
    }
    fwo_c epsc = {dd, sigma}; /* -(eps w) + i sigma */
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) {
        fwo_c t = C0;
        for (int a = 0; a < 3; a++) {
          t = cadd(t, cscale(ilam[a], nu_omega * RR[a][i][j]));
          t = cadd(t, cscale(cmul(epsc, lam[a]), MM[a][i][j]));
        }
        ae[i * 6 + j] = t;
      }
  }
  for (int i = 0; i < 6; i++) be[i] = C0;
  if (src) {
    /* tetNedelec_JoA_Complex (tetrahedral.c:794-858): exact integral
     * int N_i . sum_j lambda_j Joe_j dV with int lambda_a lambda_j dV =
     * V (1 + delta_aj) / 20; component-major accumulation as the reference */
    const fwo_c *J = (const fwo_c *)joe; /* J[vertex*3 + comp] */
    for (int i = 0; i < 6; i++) {
      int a = EDGE[i][0], b = EDGE[i][1];
      double base = ls[i] / 120.0;
      fwo_c acc = C0;
      for (int comp = 0; comp < 3; comp++)
        for (int jv = 0; jv < 4; jv++) {
          double pn = base * ((1.0 + (jv == a)) * G[b][comp] - (1.0 + (jv == b)) * G[a][comp]);
          acc = cadd(acc, cscale(J[jv * 3 + comp], pn));
        }
      /* be += (JoA - gradside * Ihe) with Ihe = 0 (make_elem_em.c:299-315) */
      be[i] = cadd(be[i], csub(acc, C0));
    }
  }
  return 0;
}

/* Delta-gap load of one element (make_elem_em.c:320-396, volt.dat dg > 0):
 * DGe_j = i dV_j per vertex, gs_ij = (grad lambda_j, N_i) =
 * (1/(6D)) (Fx_i b_j + Fy_i c_j + Fz_i d_j) with Fx_i = ls_i (yz_i - zi_i gy +
 * yi_i gz) etc. (tetNedelec_gradside / _Fxyz, tetrahedral.c:323-340,
 * 645-661), be_i = (sum_j DGe_j gs_ij) * (sigma omega) in the reference's
 * order.  dv4: the 4 vertex voltages (0 outside every gap box).          */
int fwo_dg_load(const double *v, double sigma, double omega, const double *dv4, double *be12) {
  double G[4][3], D, g[3], ls[6], e[6][3], w[6][3];
  int rc = tet_frame(v, G, &D, g, ls, e, w);
  if (rc != 0) return rc;
  const double dd = 1.0 / (6.0 * D);
  const double so = sigma * omega;
  fwo_c *be = (fwo_c *)be12;
  for (int i = 0; i < 6; i++) {
    const double Fx = ls[i] * (w[i][0] - e[i][2] * g[1] + e[i][1] * g[2]);
    const double Fy = ls[i] * (w[i][1] - e[i][0] * g[2] + e[i][2] * g[0]);
    const double Fz = ls[i] * (w[i][2] - e[i][1] * g[0] + e[i][0] * g[1]);
    fwo_c acc = C0;
    for (int j = 0; j < 4; j++) {
      const double gs = dd * (Fx * G[j][0] + Fy * G[j][1] + Fz * G[j][2]);
      acc.re += 0.0 * gs;
      acc.im += dv4[j] * gs;
    }
    be[i].re = acc.re * so;
    be[i].im = acc.im * so;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* row-list sparse matrix, lib/mymt_aij_solver/aij_matrix_func.h (TMP_AIJ)  */

/* HDDM_post_em_E (main/fullwave/fullwave.c:640-763) for one element.
 * v: 4 vertices x 3; xe12: the solution at the element's 6 edges (re, im);
 * nu: 7.957747e5 when the element's flag is a listed material, else 0
 * (:670-675).  xe is scaled by nu/omega first (:721) and the scaled values
 * feed all three outputs:
 *   rotE = (2/D) sum_k (i xe_k) ls_k e_k              (:716-729)
 *   E    = (2/D) sum_k xe_k ls_k F_k(centroid)        (:733-748)
 *   Je   = (sigma/D) sum_k xe_k ls_k F_k(centroid)    (:750-762)
 * with the geometry of tetNedelec_mk/_XYZ/_km_mk/_Fxyz
 * (lib/fem/tetrahedral.c:267-340).  Outputs 3 complex each.              */
int fwo_post_element(const double *v, const double *xe12, double nu,
                     double sigma, double omega, double *rot6, double *e6,
                     double *j6) {
  double r1[3], r2[3], r3[3];
  for (int a = 0; a < 3; a++) {
    r1[a] = v[3 + a] - v[a];
    r2[a] = v[6 + a] - v[a];
    r3[a] = v[9 + a] - v[a];
  }
  const double D = r1[0] * (r2[1] * r3[2] - r2[2] * r3[1]) +
                   r1[1] * (r2[2] * r3[0] - r2[0] * r3[2]) +
                   r1[2] * (r2[0] * r3[1] - r2[1] * r3[0]);
  if (!(D > 0.0)) return -1;
  double g[3];
  for (int a = 0; a < 3; a++) g[a] = (v[a] + v[3 + a] + v[6 + a] + v[9 + a]) * 0.25;
  const fwo_c *xe = (const fwo_c *)xe12;
  fwo_c R[3] = {C0, C0, C0}, E[3] = {C0, C0, C0};
  const double s = nu / omega;
  for (int k = 0; k < 6; k++) {
    const double *va = v + 3 * EDGE[k][0], *vb = v + 3 * EDGE[k][1];
    const double *vc = v + 3 * EDGE[k][2], *vd = v + 3 * EDGE[k][3];
    double dl[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
    double l = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    int si = 0;
    for (int a = 0; a < 3 && si == 0; a++) si = dl[a] > 0.0 ? 1 : (dl[a] < 0.0 ? -1 : 0);
    if (si == 0) return -2;
    const double ls = l * (double)si;
    double e[3], F[3];
    for (int a = 0; a < 3; a++) e[a] = vd[a] - vc[a];
    F[0] = (vc[1] * vd[2] - vd[1] * vc[2]) - e[2] * g[1] + e[1] * g[2];
    F[1] = (vc[2] * vd[0] - vd[2] * vc[0]) - e[0] * g[2] + e[2] * g[0];
    F[2] = (vc[0] * vd[1] - vd[0] * vc[1]) - e[1] * g[0] + e[0] * g[1];
    const fwo_c x = cscale(xe[k], s);
    const fwo_c ix = cmul(x, (fwo_c){0.0, 1.0});
    for (int a = 0; a < 3; a++) {
      R[a] = cadd(R[a], cscale(ix, ls * e[a]));
      E[a] = cadd(E[a], cscale(x, ls * F[a]));
    }
  }
  fwo_c *ro = (fwo_c *)rot6, *eo = (fwo_c *)e6, *jo = (fwo_c *)j6;
  for (int a = 0; a < 3; a++) {
    ro[a] = cscale(R[a], 2.0 / D);
    eo[a] = cscale(E[a], 2.0 / D);
    jo[a] = cscale(E[a], sigma / D);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
typedef struct {
  int nf;
  int *ai;      /* entries in row i                   */
  int **aij;    /* column ids j <= i                  */
  fwo_c **vij;  /* values                              */
  int *cap;
  fwo_c *b;
} tmp_t;

static void tmp_init(tmp_t *t, int nf) { /* aij_matrix_func.h:8-44 */
  t->nf = nf;
  t->ai = (int *)calloc(nf + 1, sizeof(int));
  t->aij = (int **)calloc(nf, sizeof(int *));
  t->vij = (fwo_c **)calloc(nf, sizeof(fwo_c *));
  t->cap = (int *)calloc(nf, sizeof(int));
  t->b = (fwo_c *)calloc(nf, sizeof(fwo_c));
}
static void tmp_free(tmp_t *t) {
  for (int i = 0; i < t->nf; i++) { free(t->aij[i]); free(t->vij[i]); }
  free(t->ai); free(t->aij); free(t->vij); free(t->cap); free(t->b);
  memset(t, 0, sizeof(*t));
}
static void tmp_push(tmp_t *t, int row, int col, fwo_c v) {
  if (t->ai[row] == t->cap[row]) {
    int nc = t->cap[row] ? 2 * t->cap[row] : 8;
    t->aij[row] = (int *)realloc(t->aij[row], nc * sizeof(int));
    t->vij[row] = (fwo_c *)realloc(t->vij[row], nc * sizeof(fwo_c));
    t->cap[row] = nc;
  }
  t->aij[row][t->ai[row]] = col;
  t->vij[row][t->ai[row]] = v;
  t->ai[row]++;
}
/* Copy_TMP_AIJ (aij_matrix_func.h:395-453) */
static void tmp_copy(tmp_t *d, const tmp_t *s) {
  tmp_init(d, s->nf);
  for (int i = 0; i < s->nf; i++) {
    int n = s->ai[i];
    d->ai[i] = n;
    d->cap[i] = n;
    d->aij[i] = (int *)malloc((n ? n : 1) * sizeof(int));
    d->vij[i] = (fwo_c *)malloc((n ? n : 1) * sizeof(fwo_c));
    memcpy(d->aij[i], s->aij[i], n * sizeof(int));
    memcpy(d->vij[i], s->vij[i], n * sizeof(fwo_c));
  }
  memcpy(d->b, s->b, s->nf * sizeof(fwo_c));
}
/* SetAe, symmetric storage (aij_matrix_func.h:260-323): entries with
 * column index above the row are dropped; others are added to an existing
 * entry found by linear search, or appended.                              */
static void tmp_set_ae(tmp_t *t, const int *ind, fwo_c *ae) {
  for (int col = 0; col < 6; col++)
    for (int row = 0; row < 6; row++)
      if (ind[col] < ind[row]) ae[col * 6 + row] = C0;
  for (int col = 0; col < 6; col++) {
    int icol = ind[col];
    for (int row = 0; row < 6; row++) {
      int irow = ind[row];
      fwo_c dd = ae[col * 6 + row];
      if (ciszero(dd)) continue;
      int found = 0;
      for (int k = 0; k < t->ai[icol]; k++)
        if (t->aij[icol][k] == irow) {
          t->vij[icol][k] = cadd(t->vij[icol][k], dd);
          found = 1;
          break;
        }
      if (!found) tmp_push(t, icol, irow, dd);
    }
  }
}
/* ascending column order per row (sort_ic, lib/std/sort.c:88-108) */
static void sort_pairs(int n, int *k, fwo_c *v) {
  for (int i = 1; i < n; i++) {
    int kk = k[i];
    fwo_c vv = v[i];
    int j = i - 1;
    while (j >= 0 && k[j] > kk) { k[j + 1] = k[j]; v[j + 1] = v[j]; j--; }
    k[j + 1] = kk;
    v[j + 1] = vv;
  }
}
/* the reference sorts with an O(n^2) bubble sort on every call; the cost is
 * kept so that the CPU timing reflects the reference algorithm            */
static void bubble_pairs(int n, int *k, fwo_c *v) {
  for (int i = 1; i < n; i++)
    for (int j = n - 1; j >= i; j--)
      if (k[j - 1] > k[j]) {
        int t = k[j - 1]; k[j - 1] = k[j]; k[j] = t;
        fwo_c c = v[j - 1]; v[j - 1] = v[j]; v[j] = c;
      }
}
/* DeleteZero (aij_matrix_func.h:342-391) */
static void tmp_delete_zero(tmp_t *t) {
  for (int i = 0; i < t->nf; i++) {
    int w = 0;
    for (int k = 0; k < t->ai[i]; k++)
      if (!ciszero(t->vij[i][k])) {
        t->aij[i][w] = t->aij[i][k];
        t->vij[i][w] = t->vij[i][k];
        w++;
      }
    t->ai[i] = w;
  }
}
/* SetBC, symmetric storage (aij_matrix_func.h:60-136).  ibc/vbc are
 * sorted in place (callers pass copies, as hddm_solver.h:487-491 does).   */
static void tmp_setbc(tmp_t *t, int nbc, int *ibc, fwo_c *vbc, int bubble) {
  if (bubble) {
    bubble_pairs(nbc, ibc, vbc);
    for (int i = 0; i < t->nf; i++) bubble_pairs(t->ai[i], t->aij[i], t->vij[i]);
  } else {
    sort_pairs(nbc, ibc, vbc);
    for (int i = 0; i < t->nf; i++) sort_pairs(t->ai[i], t->aij[i], t->vij[i]);
  }
  /* BC rows: move the row to the load, then make it the identity row      */
  for (int k = 0; k < nbc; k++) {
    int r = ibc[k];
    fwo_c v = vbc[k];
    for (int q = 0; q < t->ai[r]; q++) {
      int c = t->aij[r][q];
      t->b[c] = csub(t->b[c], cmul(t->vij[r][q], v));
    }
    if (t->cap[r] < 1) {
      t->aij[r] = (int *)realloc(t->aij[r], sizeof(int));
      t->vij[r] = (fwo_c *)realloc(t->vij[r], sizeof(fwo_c));
      t->cap[r] = 1;
    }
    t->aij[r][0] = r;
    t->vij[r][0] = C1;
    t->ai[r] = 1;
  }
  /* BC columns in the remaining rows (merge walk over sorted lists)       */
  for (int i = 0; i < t->nf; i++) {
    int bc = 0;
    for (int q = 0; q < t->ai[i]; q++) {
      int j = t->aij[i][q];
      if (j == i) continue;
      for (; bc < nbc; bc++) {
        if (j < ibc[bc]) break;
        if (ibc[bc] == j) {
          t->b[i] = csub(t->b[i], cmul(t->vij[i][q], vbc[bc]));
          t->vij[i][q] = C0;
          bc++;
          break;
        }
      }
    }
  }
  for (int k = 0; k < nbc; k++) t->b[ibc[k]] = vbc[k];
  tmp_delete_zero(t);
}
/* MxV_TMP_AIJ, symmetric (aij_matrix_func.h:624-658) */
static void tmp_mxv(const tmp_t *t, const fwo_c *x, fwo_c *y) {
  for (int i = 0; i < t->nf; i++) y[i] = C0;
  for (int i = 0; i < t->nf; i++)
    for (int q = 0; q < t->ai[i]; q++) {
      int j = t->aij[i][q];
      y[i] = cadd(y[i], cmul(t->vij[i][q], x[j]));
      if (j != i) y[j] = cadd(y[j], cmul(t->vij[i][q], x[i]));
    }
}

/* ------------------------------------------------------------------------ */
/* skyline LDL^T, lib/hddm/advlas_skyline.c                                 */
/* ------------------------------------------------------------------------ */
/* advlas_copy_nzindex_to_skyindex (advlas_skyline.c:12-25) */
static int sky_index(const tmp_t *t, int *idx1, int *idx2) {
  int n = t->nf;
  for (int i = 0; i < n; i++) idx1[i] = t->aij[i][0];
  idx2[0] = 0;
  for (int i = 1; i < n; i++) idx2[i] = idx2[i - 1] + i - idx1[i];
  return idx2[n - 1] + n;
}

void fwo_sky_factor(int n, const int *idx1, const int *idx2, double *matd) {
  double _Complex *S = (double _Complex *)matd;
  double _Complex *dg = (double _Complex *)malloc(n * sizeof(double _Complex));
  dg[0] = 1 / S[0];
  for (int i = 1; i < n; i++) dg[i] = S[idx2[i] + i];
  for (int i = 1; i < n; i++) {
    int fi = idx2[i];
    for (int j = idx1[i]; j < i; j++) {
      int fj = idx2[j];
      int k0 = idx1[i] > idx1[j] ? idx1[i] : idx1[j];
      double _Complex r = 0.0;
      for (int k = k0; k < j; k++) r += S[fi + k] * S[fj + k];
      S[fi + j] -= r;
    }
    double _Complex r = 0.0;
    for (int k = idx1[i]; k < i; k++) {
      r += S[fi + k] * S[fi + k] * dg[k];
      S[fi + k] *= dg[k];
    }
    dg[i] = 1 / (dg[i] - r);
  }
  for (int i = 0; i < n; i++) S[idx2[i] + i] = dg[i];
  free(dg);
}

void fwo_sky_solve(int n, const int *idx1, const int *idx2, const double *matd,
                   double *vd) {
  const double _Complex *S = (const double _Complex *)matd;
  double _Complex *v = (double _Complex *)vd;
  for (int i = 1; i < n; i++) {
    int ii = idx2[i];
    double _Complex r = 0.0;
    for (int j = idx1[i]; j < i; j++) r += S[ii + j] * v[j];
    v[i] -= r;
  }
  for (int i = 0; i < n; i++) v[i] *= S[idx2[i] + i];
  for (int i = n - 1; i >= 1; i--) {
    int ii = idx2[i];
    double _Complex r = v[i];
    for (int j = idx1[i]; j < i; j++) v[j] -= S[ii + j] * r;
  }
}

/* ------------------------------------------------------------------------ */
/* CSR (AIJ_MAT) helpers, aij_matrix_func.h:140-235, 590-682                */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nf;
  int *ai, *aij;
  fwo_c *vij;
} csr_t;

static void csr_from_tmp(csr_t *c, const tmp_t *t) { /* Assem, :140-235 */
  c->nf = t->nf;
  c->ai = (int *)malloc((t->nf + 1) * sizeof(int));
  int nnz = 0;
  for (int i = 0; i < t->nf; i++) nnz += t->ai[i];
  c->aij = (int *)malloc((nnz ? nnz : 1) * sizeof(int));
  c->vij = (fwo_c *)malloc((nnz ? nnz : 1) * sizeof(fwo_c));
  int p = 0;
  for (int i = 0; i < t->nf; i++) {
    c->ai[i] = p;
    for (int q = 0; q < t->ai[i]; q++, p++) {
      c->aij[p] = t->aij[i][q];
      c->vij[p] = t->vij[i][q];
    }
  }
  c->ai[t->nf] = p;
}
static void csr_free(csr_t *c) { free(c->ai); free(c->aij); free(c->vij); }

/* symmetric CSR MxV (aij_matrix_func.h:590-620) */
static void csr_mxv(int m, const int *ai, const int *aij, const fwo_c *vij,
                    const fwo_c *vec, fwo_c *x) {
  for (int i = 0; i < m; i++) x[i] = C0;
  for (int i = 0; i < m; i++) {
    int j;
    for (j = ai[i]; j < ai[i + 1]; j++) {
      int jj = aij[j];
      fwo_c dd = vij[j];
      x[i] = cadd(x[i], cmul(dd, vec[jj]));
      x[jj] = cadd(x[jj], cmul(dd, vec[i]));
    }
    j--;
    if (j >= ai[i] && i == aij[j]) x[i] = csub(x[i], cmul(vij[j], vec[aij[j]]));
  }
}
/* getAIJ, symmetric (aij_matrix_func.h:662-682): lower entry (i, j) */
static fwo_c csr_get(int i, int j, const int *ai, const int *aij, const fwo_c *v) {
  if (i < j) { int t = i; i = j; j = t; }
  for (int k = ai[i]; k < ai[i + 1]; k++)
    if (aij[k] == j) return v[k];
  return C0;
}

/* ICC make (aij_pc_func.h:36-108); pc = [L (nnz) | d (n)] */
static void icc_make(int nf, const int *ai, const int *aij, const fwo_c *vij,
                     double af, fwo_c *M) {
  fwo_c *L = M, *d = M + ai[nf];
  if (af > 3.0 || af < 1.0) af = 1.05;
  for (int col = 0; col < nf; col++) {
    int ii = ai[col], jj = ai[col + 1] - 1;
    fwo_c dd;
    for (int row = ii; row < jj; row++) {
      dd = C0;
      for (int i = ii; i < row; i++)
        dd = cadd(dd, cmul(d[aij[i]], cmul(csr_get(col, aij[i], ai, aij, L),
                                           csr_get(aij[row], aij[i], ai, aij, L))));
      L[row] = csub(vij[row], dd);
    }
    dd = C0;
    for (int i = ii; i < jj; i++) {
      fwo_c q = csr_get(col, aij[i], ai, aij, L);
      dd = cadd(dd, cmul(d[aij[i]], cmul(q, q)));
    }
    L[jj] = csub(cscale(vij[jj], af), dd);
    d[col] = cdiv(C1, L[jj]);
  }
}
/* ICC set (aij_pc_func.h:112-167) */
static void icc_set(int nf, const int *ai, const int *aij, const fwo_c *M,
                    fwo_c *x, const fwo_c *b) {
  const fwo_c *L = M, *d = M + ai[nf];
  for (int col = 0; col < nf; col++) {
    int ii = ai[col], jj = ai[col + 1] - 1;
    fwo_c dd = C0;
    for (int row = ii; row < jj; row++) dd = cadd(dd, cmul(x[aij[row]], L[row]));
    x[col] = cdiv(csub(b[col], dd), L[jj]);
  }
  for (int col = nf - 1; col >= 0; col--) {
    int ii = ai[col], jj = ai[col + 1] - 2;
    for (int row = jj; row >= ii; row--)
      x[aij[row]] = csub(x[aij[row]], cmul(x[col], cmul(L[row], d[aij[row]])));
  }
}

void fwo_icc_make(int n, const int *rowptr, const int *col, const double *val,
                  double ic_shift, double *pc) {
  icc_make(n, rowptr, col, (const fwo_c *)val, ic_shift, (fwo_c *)pc);
}

/* COCG with scaled directions (aij_solver_cg.h:113-355).  status codes
 * follow lib/std/defs.h:164-169.                                          */
#define ST_DIVERGED (-2)
#define ST_MAXITE (-3)
static int cocg_run(int nf, const int *ai, const int *aij, const fwo_c *vij,
                    const fwo_c *b, fwo_c *x, double eps, double div,
                    const fwo_c *pc, double *hist) {
  fwo_c *r = (fwo_c *)malloc(nf * sizeof(fwo_c));
  fwo_c *v = (fwo_c *)malloc(nf * sizeof(fwo_c));
  fwo_c *aq = (fwo_c *)malloc(nf * sizeof(fwo_c));
  fwo_c *q = (fwo_c *)malloc(nf * sizeof(fwo_c));
  int max_ite = nf * 10, ret = -1, n;
  csr_mxv(nf, ai, aij, vij, x, aq);
  for (int i = 0; i < nf; i++) r[i] = csub(b[i], aq[i]);
  double bn = sqrt(cabs_(cinner(nf, b, b)));
  if (bn < NEARLY_ZERO) { ret = 0; goto END; }
  bn = 1.0 / bn;
  double res = sqrt(cabs_(cinner(nf, r, r))) * bn;
  if (hist) hist[0] = res;
  if (res <= eps) { ret = 0; goto END; }
  if (pc) icc_set(nf, ai, aij, pc, v, r); else memcpy(v, r, nf * sizeof(fwo_c));
  fwo_c rv = crecip(cbilinear(nf, v, r));
  for (int i = 0; i < nf; i++) q[i] = cmul(v[i], rv);
  for (n = 1; n < max_ite; n++) {
    csr_mxv(nf, ai, aij, vij, q, aq);
    fwo_c alpha = cbilinear(nf, q, aq);
    if (ciszero(alpha)) { ret = -5; goto END; } /* the reference exits the process here */
    alpha = cdiv(C1, alpha);
    for (int i = 0; i < nf; i++) x[i] = cadd(x[i], cmul(alpha, q[i]));
    alpha = csub(C0, alpha);
    for (int i = 0; i < nf; i++) r[i] = cadd(r[i], cmul(alpha, aq[i]));
    res = sqrt(cabs_(cinner(nf, r, r))) * bn;
    if (hist) hist[n] = res;
    if (res <= eps) { ret = n; goto END; }
    if (res >= div) { ret = ST_DIVERGED; goto END; }
    if (pc) icc_set(nf, ai, aij, pc, v, r); else memcpy(v, r, nf * sizeof(fwo_c));
    rv = crecip(cbilinear(nf, v, r));
    for (int i = 0; i < nf; i++) q[i] = cadd(q[i], cmul(v[i], rv));
  }
  ret = ST_MAXITE;
END:
  free(r); free(v); free(aq); free(q);
  return ret;
}

int fwo_cocg_icc(int n, const int *rowptr, const int *col, const double *val,
                 const double *b, double *x, double eps, double ic_shift,
                 int use_icc, double *hist) {
  fwo_c *pc = NULL;
  if (use_icc) {
    pc = (fwo_c *)malloc((rowptr[n] + n) * sizeof(fwo_c));
    icc_make(n, rowptr, col, (const fwo_c *)val, ic_shift, pc);
  }
  int st = cocg_run(n, rowptr, col, (const fwo_c *)val, (const fwo_c *)b,
                    (fwo_c *)x, eps, 1.0e+10, pc, hist);
  free(pc);
  return st;
}

/* ------------------------------------------------------------------------ */
/* per-subdomain model                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int part;
  tmp_t A;          /* assembled, sorted, zero-free lower triangle + f     */
  int nbc, ninbc;
  int *ibc;         /* [essential | interface nfids]                      */
  fwo_c *vbc;
  const int *nfid, *ln;
  int nlen;
  int *idx1, *idx2;
  double *fac;      /* 2*nlen                                              */
  fwo_c *pcd;       /* diag(A) for the interface preconditioner           */
  fwo_c *icc;       /* COCG mode: ICC factor of the Dirichlet matrix       */
  csr_t Ab;         /* COCG mode: Dirichlet matrix as CSR                  */
  fwo_c *x;         /* last solution                                       */
} sub_t;

struct fwo_model {
  int nsub, npart;
  sub_t *s;
};

/* HDDM_make_TMP_AIJ_em_E (main/fullwave/fullwave.c:221-259) */
static int assemble_sub(const fwh_mesh_desc *m, int s, tmp_t *A) {
  int nf = m->nmp[s];
  tmp_init(A, nf);
  const double *crd = m->crd + 3 * (size_t)m->ap_off[s];
  const double *joe = m->joe ? m->joe + 6 * (size_t)m->ap_off[s] : NULL;
  for (int el = m->el_off[s]; el < m->el_off[s + 1]; el++) {
    const int *ne = m->nop + 10 * (size_t)el;
    double xyz[12], J[24];
    for (int i = 0; i < 4; i++)
      for (int a = 0; a < 3; a++) xyz[3 * i + a] = crd[3 * ne[i] + a];
    int src = m->src ? m->src[el] : 0;
    if (src && joe)
      for (int i = 0; i < 4; i++) memcpy(J + 6 * i, joe + 6 * ne[i], 6 * sizeof(double));
    else
      memset(J, 0, sizeof(J));
    double ae[72], be[12];
    if (fwo_element(xyz, m->eps[el], m->sigma[el], m->omega,
                    m->pml ? m->pml + 6 * (size_t)el : NULL, m->dgv ? 0 : src, J, ae, be) != 0)
      return -1;
    if (m->dgv) {
      /* delta-gap mode: the gap load replaces the current source
       * (make_elem_em.c:274-279 vs :320-396)                             */
      const double *dv = m->dgv + (size_t)m->ap_off[s];
      double d4[4] = {dv[ne[0]], dv[ne[1]], dv[ne[2]], dv[ne[3]]};
      if (fwo_dg_load(xyz, m->sigma[el], m->omega, d4, be) != 0) return -1;
    }
    int ind[6];
    for (int k = 0; k < 6; k++) ind[k] = ne[4 + k];
    tmp_set_ae(A, ind, (fwo_c *)ae);
    for (int k = 0; k < 6; k++) A->b[ind[k]] = cadd(A->b[ind[k]], ((fwo_c *)be)[k]);
  }
  for (int i = 0; i < nf; i++) sort_pairs(A->ai[i], A->aij[i], A->vij[i]);
  tmp_delete_zero(A);
  return 0;
}

static void sub_free(sub_t *u) {
  tmp_free(&u->A);
  free(u->ibc); free(u->vbc); free(u->idx1); free(u->idx2); free(u->fac);
  free(u->pcd); free(u->icc); free(u->x);
  if (u->Ab.ai) csr_free(&u->Ab);
}

static int sub_setup(const fwh_mesh_desc *m, int s, sub_t *u) {
  memset(u, 0, sizeof(*u));
  u->part = m->sub_part[s];
  if (assemble_sub(m, s, &u->A) != 0) return -1;
  int ne = m->bc_off[s + 1] - m->bc_off[s];
  u->ninbc = m->ib_off[s + 1] - m->ib_off[s];
  u->nbc = ne + u->ninbc;
  u->ibc = (int *)malloc((u->nbc + 1) * sizeof(int));
  u->vbc = (fwo_c *)calloc(u->nbc + 1, sizeof(fwo_c)); /* init_bc: zeros */
  memcpy(u->ibc, m->ibc + m->bc_off[s], ne * sizeof(int));
  memcpy(u->ibc + ne, m->ib_nfid + m->ib_off[s], u->ninbc * sizeof(int));
  u->nfid = m->ib_nfid + m->ib_off[s];
  u->ln = m->ib_ln + m->ib_off[s];
  int n = u->A.nf;
  u->x = (fwo_c *)calloc(n, sizeof(fwo_c));
  /* diag(A) for HDDM_PC_DIAG (lib/hddm/hddm_pc.h:52-86) */
  u->pcd = (fwo_c *)calloc(n, sizeof(fwo_c));
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < u->A.ai[i]; q++)
      if (u->A.aij[i][q] == i) { u->pcd[i] = ciszero(u->A.vij[i][q]) ? C1 : u->A.vij[i][q]; break; }
    if (ciszero(u->pcd[i])) u->pcd[i] = C1;
  }
  return 0;
}

/* Dirichlet copy with the current vbc: copy A, optionally zero the load,
 * SetBC on copies of ibc/vbc (hddm_solver.h:485-493, 581-592).            */
static void sub_dirichlet(sub_t *u, tmp_t *T, int zero_load, int bubble) {
  tmp_copy(T, &u->A);
  if (zero_load) memset(T->b, 0, T->nf * sizeof(fwo_c));
  int *ib = (int *)malloc((u->nbc + 1) * sizeof(int));
  fwo_c *vb = (fwo_c *)malloc((u->nbc + 1) * sizeof(fwo_c));
  memcpy(ib, u->ibc, u->nbc * sizeof(int));
  memcpy(vb, u->vbc, u->nbc * sizeof(fwo_c));
  tmp_setbc(T, u->nbc, ib, vb, bubble);
  free(ib); free(vb);
}

static void sub_factor(sub_t *u, int bubble) { /* hddm_solver.h:495-514 */
  tmp_t T;
  sub_dirichlet(u, &T, 0, bubble);
  int n = T.nf;
  u->idx1 = (int *)malloc((n + 1) * sizeof(int));
  u->idx2 = (int *)malloc((n + 1) * sizeof(int));
  u->nlen = sky_index(&T, u->idx1, u->idx2);
  u->fac = (double *)calloc(2 * (size_t)u->nlen, sizeof(double));
  fwo_c *F = (fwo_c *)u->fac;
  for (int i = 0; i < n; i++) /* advlas_copy_nzmat_to_skymat :49-60 */
    for (int q = 0; q < T.ai[i]; q++) F[u->idx2[i] + T.aij[i][q]] = T.vij[i][q];
  fwo_sky_factor(n, u->idx1, u->idx2, u->fac);
  tmp_free(&T);
}

fwo_model *fwo_model_build(const fwh_mesh_desc *m) {
  fwo_model *md = (fwo_model *)calloc(1, sizeof(fwo_model));
  md->nsub = m->nsub;
  md->npart = m->npart;
  md->s = (sub_t *)calloc(m->nsub, sizeof(sub_t));
  for (int s = 0; s < m->nsub; s++) {
    if (sub_setup(m, s, &md->s[s]) != 0) { fwo_model_free(md); return NULL; }
    sub_factor(&md->s[s], 0);
  }
  return md;
}

void fwo_model_free(fwo_model *md) {
  if (!md) return;
  for (int s = 0; s < md->nsub; s++) sub_free(&md->s[s]);
  free(md->s);
  free(md);
}

void fwo_model_sizes(const fwo_model *md, int s, int *n, int *nlen, int *nnz, int *nbc) {
  const sub_t *u = &md->s[s];
  int z = 0;
  for (int i = 0; i < u->A.nf; i++) z += u->A.ai[i];
  if (n) *n = u->A.nf;
  if (nlen) *nlen = u->nlen;
  if (nnz) *nnz = z;
  if (nbc) *nbc = u->nbc;
}

void fwo_model_get(const fwo_model *md, int s, int *rowptr, int *col, double *val,
                   double *f, int *idx1, int *idx2, double *factor) {
  const sub_t *u = &md->s[s];
  int p = 0, n = u->A.nf;
  for (int i = 0; i < n; i++) {
    if (rowptr) rowptr[i] = p;
    for (int q = 0; q < u->A.ai[i]; q++, p++) {
      if (col) col[p] = u->A.aij[i][q];
      if (val) { val[2 * p] = u->A.vij[i][q].re; val[2 * p + 1] = u->A.vij[i][q].im; }
    }
  }
  if (rowptr) rowptr[n] = p;
  if (f) memcpy(f, u->A.b, n * sizeof(fwo_c));
  if (idx1) memcpy(idx1, u->idx1, n * sizeof(int));
  if (idx2) memcpy(idx2, u->idx2, n * sizeof(int));
  if (factor) memcpy(factor, u->fac, 2 * (size_t)u->nlen * sizeof(double));
}

/* hddm_solver.h:559-622: Dirichlet solve with zero load, then y = A x */
static void sub_solve_dom(sub_t *u, int zero_load, int bubble, fwo_c *y) {
  tmp_t T;
  sub_dirichlet(u, &T, zero_load, bubble);
  int n = T.nf;
  memcpy(u->x, T.b, n * sizeof(fwo_c));
  tmp_free(&T);
  fwo_sky_solve(n, u->idx1, u->idx2, u->fac, (double *)u->x);
  tmp_mxv(&u->A, u->x, y);
}

void fwo_model_schur_sub(fwo_model *md, int s, const double *vbc, double *x, double *y) {
  sub_t *u = &md->s[s];
  int ne = u->nbc - u->ninbc;
  for (int i = 0; i < ne; i++) u->vbc[i] = C0;
  memcpy(u->vbc + ne, vbc, u->ninbc * sizeof(fwo_c));
  sub_solve_dom(u, 1, 0, (fwo_c *)y);
  if (x) memcpy(x, u->x, u->A.nf * sizeof(fwo_c));
}

/* ------------------------------------------------------------------------ */
/* s-mode HDDM, lib/hddm/hddm_solver.h:59-183, hddm_solver_s.h:7-438,       */
/* hddm_solver_sp.h:8-69, hddm_pc.h:52-137                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int infree, outfree, tfree;
  fwo_c *P, *Q, *R, *Z, *U, *PC;
  fwo_c rz;
  double r0;
} hvec_t;

typedef struct {
  const fwh_mesh_desc *m;
  fwo_model *md;
  hvec_t *hv;
  fwo_opts o;
  long sub_solves;
} hddm_t;

/* __HDDM_SOLVER_ITERATION_S_GATHER__ (hddm_solver_s.h:404-424) on any
 * part vector (Q or PC)                                                    */
static void s_gather(hddm_t *H, int which) {
  const fwh_mesh_desc *m = H->m;
  int np = m->npart;
  for (int p = 0; p < np; p++)
    for (int q = p + 1; q < np; q++) {
      int a0 = m->op_off[p * np + q], a1 = m->op_off[p * np + q + 1];
      int b0 = m->op_off[q * np + p];
      fwo_c *dst = which ? H->hv[p].PC : H->hv[p].Q;
      fwo_c *src = which ? H->hv[q].PC : H->hv[q].Q;
      for (int j = 0; j < a1 - a0; j++) {
        int ii = m->op_ibid[a0 + j], jj = m->op_ibid[b0 + j];
        dst[ii] = cadd(dst[ii], src[jj]);
      }
    }
}
/* owner -> sharer copy of P, then Q = 0 (hddm_solver_s.h:237-252) */
static void s_bcast_p(hddm_t *H) {
  const fwh_mesh_desc *m = H->m;
  int np = m->npart;
  for (int p = 0; p < np; p++)
    for (int q = p + 1; q < np; q++) {
      int a0 = m->op_off[p * np + q], a1 = m->op_off[p * np + q + 1];
      int b0 = m->op_off[q * np + p];
      for (int j = 0; j < a1 - a0; j++) {
        int ii = m->op_ibid[b0 + j], jj = m->op_ibid[a0 + j];
        H->hv[q].P[ii] = H->hv[p].P[jj];
      }
    }
  for (int p = 0; p < np; p++) memset(H->hv[p].Q, 0, H->hv[p].tfree * sizeof(fwo_c));
}

/* subdomain solve of one loop kind: 0 = first (hddm_solver.h:461-555),
 * 1 = iteration (:559-622), 2 = last (:626-693)                            */
static void s_solve(hddm_t *H, int s, int kind) {
  sub_t *u = &H->md->s[s];
  hvec_t *hv = &H->hv[u->part];
  int n = u->A.nf;
  fwo_c *y = (fwo_c *)malloc(n * sizeof(fwo_c));
  if (kind > 0) /* renewal_bc_sp (hddm_solver_sp.h:58-69) */
    for (int i = 0, j = u->nbc - u->ninbc; i < u->ninbc; i++, j++) u->vbc[j] = hv->P[u->ln[i]];
  if (H->o.dom_solver == 0) {
    sub_solve_dom(u, kind == 1, 1, y);
  } else {
    /* COCG + ICC subdomain solver (hddm_solver.h:297-457) */
    tmp_t T;
    sub_dirichlet(u, &T, kind == 1, 1);
    if (!u->Ab.ai) {
      csr_from_tmp(&u->Ab, &T);
      u->icc = (fwo_c *)malloc((u->Ab.ai[n] + n) * sizeof(fwo_c));
      icc_make(n, u->Ab.ai, u->Ab.aij, u->Ab.vij, H->o.ic_shift, u->icc);
    }
    memset(u->x, 0, n * sizeof(fwo_c));
    cocg_run(n, u->Ab.ai, u->Ab.aij, u->Ab.vij, T.b, u->x, H->o.sol_eps, 1.0e+10, u->icc, NULL);
    tmp_free(&T);
    tmp_mxv(&u->A, u->x, y);
  }
  if (kind != 1)
    for (int i = 0; i < n; i++) y[i] = csub(y[i], u->A.b[i]);
  /* COM_res_sp(_1st) (hddm_solver_sp.h:8-54) */
  for (int i = 0; i < u->ninbc; i++) hv->Q[u->ln[i]] = cadd(hv->Q[u->ln[i]], y[u->nfid[i]]);
  if (kind == 0)
    for (int i = 0; i < u->ninbc; i++) hv->PC[u->ln[i]] = cadd(hv->PC[u->ln[i]], u->pcd[u->nfid[i]]);
  free(y);
  H->sub_solves++;
}

/* __HDDM_SOLVER_ITERATION_S_1ST__ (hddm_solver_s.h:158-256) */
static int s_iteration_1st(hddm_t *H, double *resid) {
  int np = H->m->npart;
  s_gather(H, 0);
  for (int p = 0; p < np; p++)
    memcpy(H->hv[p].R, H->hv[p].Q + H->hv[p].outfree, H->hv[p].infree * sizeof(fwo_c));
  /* hddm_pc_CS_DIAG_gather_s (hddm_pc.h:108-137): per part, gather from
   * higher parts, then invert                                               */
  {
    const fwh_mesh_desc *m = H->m;
    for (int p = 0; p < np; p++) {
      for (int q = p + 1; q < np; q++) {
        int a0 = m->op_off[p * np + q], a1 = m->op_off[p * np + q + 1];
        int b0 = m->op_off[q * np + p];
        for (int j = 0; j < a1 - a0; j++) {
          int ii = m->op_ibid[a0 + j], jj = m->op_ibid[b0 + j];
          H->hv[p].PC[ii] = cadd(H->hv[p].PC[ii], H->hv[q].PC[jj]);
        }
      }
      for (int i = 0; i < H->hv[p].tfree; i++) {
        fwo_c t = H->hv[p].PC[i];
        H->hv[p].PC[i] = ciszero(t) ? C1 : cdiv(C1, t);
      }
    }
  }
  double resi = 0.0;
  fwo_c rz = C0;
  for (int p = 0; p < np; p++) {
    hvec_t *h = &H->hv[p];
    for (int i = 0; i < h->infree; i++) h->Z[i] = cmul(h->R[i], h->PC[h->outfree + i]);
    memcpy(h->P + h->outfree, h->Z, h->infree * sizeof(fwo_c));
    for (int i = 0; i < h->infree; i++) {
      double t = cabs_(h->R[i]);
      if (t > resi) resi = t;
    }
  }
  int status = resi < NEARLY_ZERO ? 2 : 1;
  if (resid) resid[0] = 1.0;
  double r0 = resi;
  resi = 1.0 / resi;
  for (int p = 0; p < np; p++)
    for (int i = 0; i < H->hv[p].infree; i++) rz = cadd(rz, cmul(H->hv[p].R[i], H->hv[p].Z[i]));
  for (int p = 0; p < np; p++) { H->hv[p].r0 = resi; H->hv[p].rz = rz; }
  s_bcast_p(H);
  H->hv[0].r0 = resi;
  (void)r0;
  return status;
}

/* __HDDM_SOLVER_ITERATION_S__ (hddm_solver_s.h:260-400) */
static int s_iteration(hddm_t *H, int loop, double *resid) {
  int np = H->m->npart;
  s_gather(H, 0);
  fwo_c rz = H->hv[0].rz, pq = C0, malpha, beta;
  double resi = 0.0;
  for (int p = 0; p < np; p++) {
    hvec_t *h = &H->hv[p];
    if (H->o.perturb)
      for (int i = h->infree - 1; i >= 0; i--) pq = cadd(pq, cmul(h->P[h->outfree + i], h->Q[h->outfree + i]));
    else
      for (int i = 0, j = h->outfree; i < h->infree; i++, j++) pq = cadd(pq, cmul(h->P[j], h->Q[j]));
  }
  if (cabs_(rz) < NEARLY_ZERO && cabs_(pq) < NEARLY_ZERO) malpha = C0;
  else malpha = csub(C0, cdiv(rz, pq));
  for (int p = 0; p < np; p++) {
    hvec_t *h = &H->hv[p];
    for (int i = 0, j = h->outfree; i < h->infree; i++, j++) {
      h->U[i] = cadd(h->U[i], cmul(malpha, h->P[j]));
      h->R[i] = cadd(h->R[i], cmul(malpha, h->Q[j]));
      double t = cabs_(h->R[i]);
      if (t > resi) resi = t;
    }
  }
  resi *= H->hv[0].r0;
  if (resid) resid[loop] = resi;
  for (int p = 0; p < np; p++) { /* hddm_pc_CS_DIAG_set (hddm_pc.h:90-103) */
    hvec_t *h = &H->hv[p];
    for (int i = 0, j = h->outfree; i < h->infree; i++, j++) h->Z[i] = cmul(h->R[i], h->PC[j]);
  }
  rz = C0;
  for (int p = 0; p < np; p++)
    if (H->o.perturb)
      for (int i = H->hv[p].infree - 1; i >= 0; i--) rz = cadd(rz, cmul(H->hv[p].R[i], H->hv[p].Z[i]));
    else
      for (int i = 0; i < H->hv[p].infree; i++) rz = cadd(rz, cmul(H->hv[p].R[i], H->hv[p].Z[i]));
  if (cabs_(rz) < NEARLY_ZERO && cabs_(H->hv[0].rz) < NEARLY_ZERO) beta = C0;
  else beta = cdiv(rz, H->hv[0].rz);
  for (int p = 0; p < np; p++) H->hv[p].rz = rz;
  for (int p = 0; p < np; p++) {
    hvec_t *h = &H->hv[p];
    for (int i = 0, j = h->outfree; i < h->infree; i++, j++) h->P[j] = cadd(h->Z[i], cmul(beta, h->P[j]));
  }
  int status = 1;
  if (resi < H->o.hddm_eps) {
    status = 2;
    for (int p = 0; p < np; p++)
      memcpy(H->hv[p].P + H->hv[p].outfree, H->hv[p].U, H->hv[p].infree * sizeof(fwo_c));
  }
  s_bcast_p(H);
  return status;
}

static void hddm_init(hddm_t *H, const fwh_mesh_desc *m, fwo_model *md, const fwo_opts *o) {
  memset(H, 0, sizeof(*H));
  H->m = m;
  H->md = md;
  H->o = *o;
  H->hv = (hvec_t *)calloc(m->npart, sizeof(hvec_t));
  for (int p = 0; p < m->npart; p++) { /* make_HDDM_VEC (hddm_solver.h:273-293) */
    hvec_t *h = &H->hv[p];
    h->infree = m->infree[p];
    h->outfree = m->outfree[p];
    h->tfree = h->infree + h->outfree;
    h->P = (fwo_c *)calloc(h->tfree + 1, sizeof(fwo_c));
    h->Q = (fwo_c *)calloc(h->tfree + 1, sizeof(fwo_c));
    h->PC = (fwo_c *)calloc(h->tfree + 1, sizeof(fwo_c));
    h->R = (fwo_c *)calloc(h->infree + 1, sizeof(fwo_c));
    h->Z = (fwo_c *)calloc(h->infree + 1, sizeof(fwo_c));
    h->U = (fwo_c *)calloc(h->infree + 1, sizeof(fwo_c));
  }
}
static void hddm_free(hddm_t *H) {
  for (int p = 0; p < H->m->npart; p++) {
    hvec_t *h = &H->hv[p];
    free(h->P); free(h->Q); free(h->PC); free(h->R); free(h->Z); free(h->U);
  }
  free(H->hv);
}

static fwo_model *model_for_hddm(const fwh_mesh_desc *m) {
  fwo_model *md = (fwo_model *)calloc(1, sizeof(fwo_model));
  md->nsub = m->nsub;
  md->npart = m->npart;
  md->s = (sub_t *)calloc(m->nsub, sizeof(sub_t));
  for (int s = 0; s < m->nsub; s++)
    if (sub_setup(m, s, &md->s[s]) != 0) { fwo_model_free(md); return NULL; }
  return md;
}

static int hddm_solve_impl(const fwh_mesh_desc *m, const fwo_opts *o, const double *load,
                           double *resid, double *x_out, double *u_out, fwo_stats *st);

int fwo_hddm_solve(const fwh_mesh_desc *m, const fwo_opts *o, double *resid,
                   double *x_out, double *u_out, fwo_stats *st) {
  return hddm_solve_impl(m, o, NULL, resid, x_out, u_out, st);
}

int fwo_hddm_solve_load(const fwh_mesh_desc *m, const fwo_opts *o, const double *load,
                        double *resid, double *x_out, double *u_out, fwo_stats *st) {
  return hddm_solve_impl(m, o, load, resid, x_out, u_out, st);
}

static int hddm_solve_impl(const fwh_mesh_desc *m, const fwo_opts *o, const double *load,
                           double *resid, double *x_out, double *u_out, fwo_stats *st) {
  fwo_model *md = model_for_hddm(m);
  if (!md) return -1;
  if (load) { /* the per-DOF load replaces every subdomain's assembled one */
    size_t off = 0;
    for (int s = 0; s < md->nsub; s++) {
      memcpy(md->s[s].A.b, load + off, md->s[s].A.nf * sizeof(fwo_c));
      off += 2 * (size_t)md->s[s].A.nf;
    }
  }
  if (o->perturb >= 2) {
    /* rounding-level perturbation of the load (a few ulps, seeded): used
     * only to measure how the iteration count reacts to rounding          */
    unsigned long long st_ = 0x9E3779B97F4A7C15ULL * (unsigned long long)o->perturb;
    for (int s = 0; s < md->nsub; s++)
      for (int i = 0; i < md->s[s].A.nf; i++) {
        st_ = st_ * 6364136223846793005ULL + 1442695040888963407ULL;
        double r = ((double)(st_ >> 11) / 9007199254740992.0) * 2.0 - 1.0;
        md->s[s].A.b[i].re *= 1.0 + r * 0x1p-50;
        md->s[s].A.b[i].im *= 1.0 - r * 0x1p-50;
      }
  }
  hddm_t H;
  hddm_init(&H, m, md, o);
  fwo_stats S;
  memset(&S, 0, sizeof(S));
  int loop = 0, status;
  if (o->hddm_max_loop == 0) { /* hddm_solver.h:103-105: not executed */
    hddm_free(&H); fwo_model_free(md);
    if (st) *st = S;
    return 0;
  }
  /* LOOP_1ST (hddm_solver_s.h:7-52): assemble happens in model_for_hddm;
   * the first solve factors the Dirichlet matrix                          */
  for (int s = 0; s < m->nsub; s++) {
    if (o->dom_solver == 0) sub_factor(&md->s[s], 1);
    s_solve(&H, s, 0);
  }
  status = s_iteration_1st(&H, resid);
  S.r0 = 1.0 / H.hv[0].r0;
  loop++;
  double t0 = now_s();
  /* with nthreads > 1 the main loop's subdomain Schur steps run in parallel
   * into per-subdomain buffers and are accumulated afterwards in the
   * reference's order: the same sums as s_solve(H, s, 1) in turn          */
  const int par = o->nthreads > 1 && o->dom_solver == 0;
  fwo_c **ybuf = NULL;
  if (par) {
    ybuf = (fwo_c **)calloc(m->nsub, sizeof(fwo_c *));
    for (int s = 0; s < m->nsub; s++) ybuf[s] = (fwo_c *)malloc(md->s[s].A.nf * sizeof(fwo_c));
#ifdef _OPENMP
    omp_set_num_threads(o->nthreads);
#endif
  }
  while (1) { /* hddm_solver.h:125-143 */
    if (status == 2) break;
    if (par) {
      const int ns = m->nsub;
#pragma omp parallel for schedule(dynamic, 1)
      for (int s = 0; s < ns; s++) {
        sub_t *u = &md->s[s];
        hvec_t *hv = &H.hv[u->part];
        for (int i = 0, j = u->nbc - u->ninbc; i < u->ninbc; i++, j++) u->vbc[j] = hv->P[u->ln[i]];
        sub_solve_dom(u, 1, 1, ybuf[s]);
      }
      for (int s = 0; s < ns; s++) {
        sub_t *u = &md->s[s];
        hvec_t *hv = &H.hv[u->part];
        for (int i = 0; i < u->ninbc; i++) hv->Q[u->ln[i]] = cadd(hv->Q[u->ln[i]], ybuf[s][u->nfid[i]]);
        H.sub_solves++;
      }
    } else {
      for (int s = 0; s < m->nsub; s++) s_solve(&H, s, 1);
    }
    status = s_iteration(&H, loop, resid);
    S.iters++;
    if (status == 2) S.converged = 1;
    if (loop == o->hddm_max_loop) status = 2;
    loop++;
    if (o->max_sub > 0 && H.sub_solves >= o->max_sub) break;
  }
  S.seconds_loop = now_s() - t0;
  S.loops = loop;
  if (ybuf) {
    for (int s = 0; s < m->nsub; s++) free(ybuf[s]);
    free(ybuf);
  }
  for (int s = 0; s < m->nsub; s++) s_solve(&H, s, 2); /* LOOP_LAST */
  S.sub_solves = H.sub_solves;
  if (x_out) {
    size_t off = 0;
    for (int s = 0; s < m->nsub; s++) {
      memcpy(x_out + off, md->s[s].x, md->s[s].A.nf * sizeof(fwo_c));
      off += 2 * (size_t)md->s[s].A.nf;
    }
  }
  if (u_out) {
    size_t off = 0;
    for (int p = 0; p < m->npart; p++) {
      memcpy(u_out + off, H.hv[p].U, H.hv[p].infree * sizeof(fwo_c));
      off += 2 * (size_t)H.hv[p].infree;
    }
  }
  if (st) *st = S;
  hddm_free(&H);
  fwo_model_free(md);
  return 0;
}

double fwo_hddm_time_iterations(const fwh_mesh_desc *m, int n_iter, long *sub_solves, int nthreads) {
  fwo_opts o;
  memset(&o, 0, sizeof(o));
  o.hddm_eps = 0.0; /* never converges inside the window */
  o.hddm_max_loop = n_iter;
  fwo_model *md = model_for_hddm(m);
  if (!md) return -1.0;
  hddm_t H;
  hddm_init(&H, m, md, &o);
  for (int s = 0; s < m->nsub; s++) { sub_factor(&md->s[s], 1); s_solve(&H, s, 0); }
  s_iteration_1st(&H, NULL);
  long before = H.sub_solves;
  /* the subdomain Schur steps of an iteration are independent: with
   * nthreads > 1 they run in parallel into per-subdomain buffers and are
   * accumulated afterwards in the reference's order (same sums as s_solve) */
  fwo_c **ybuf = (fwo_c **)calloc(m->nsub, sizeof(fwo_c *));
  for (int s = 0; s < m->nsub; s++) ybuf[s] = (fwo_c *)malloc(md->s[s].A.nf * sizeof(fwo_c));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  double t0 = now_s();
  for (int it = 1; it <= n_iter; it++) {
    const int ns = m->nsub;
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < ns; s++) {
      sub_t *u = &md->s[s];
      hvec_t *hv = &H.hv[u->part];
      for (int i = 0, j = u->nbc - u->ninbc; i < u->ninbc; i++, j++) u->vbc[j] = hv->P[u->ln[i]];
      sub_solve_dom(u, 1, 1, ybuf[s]);
    }
    for (int s = 0; s < ns; s++) {
      sub_t *u = &md->s[s];
      hvec_t *hv = &H.hv[u->part];
      for (int i = 0; i < u->ninbc; i++) hv->Q[u->ln[i]] = cadd(hv->Q[u->ln[i]], ybuf[s][u->nfid[i]]);
      H.sub_solves++;
    }
    s_iteration(&H, it, NULL);
  }
  double dt = now_s() - t0;
  for (int s = 0; s < m->nsub; s++) free(ybuf[s]);
  free(ybuf);
  if (sub_solves) *sub_solves = H.sub_solves - before;
  hddm_free(&H);
  fwo_model_free(md);
  return dt;
}
