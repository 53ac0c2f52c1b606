// fw_kernels.hip - HIP kernels of libfwhip for gfx950 (CDNA4, wave64).
//
// Hot path of ADVENTURE_FullWave's HDDM full-wave solve:
//   * element kernel          tet_MakeElement_EM_E   lib/fem/make_elem_em.c:14-400
//   * assembly gather         mymt_CS_AIJ_SetAe      lib/mymt_aij_solver/aij_matrix_func.h:260-323
//   * Dirichlet profile       SetBC + skyindex       aij_matrix_func.h:60-136, advlas_skyline.c:12-60
//   * skyline LDL^T factor    advlas_skyline_ldl_decomposite  lib/hddm/advlas_skyline.c:79-128
//   * skyline solve + Schur   advlas_skyline_ldl_solve :134-167, hddm_solver.h:559-622
//   * interface COCG          hddm_solver_s.h:158-400, hddm_pc.h:52-137
#include "fw_kernels.h"
#include "fw_device.h"

namespace fw {

constexpr double NU = 7.957747e+05;      // lib/fem/make_elem_em.c:37

// edge k: vertices EDGE[k][0] -> EDGE[k][1]; opposite edge EDGE[k][2] -> EDGE[k][3]
// (the reference's tet_n table, lib/fem/tetrahedral.h:19-24)
__constant__ int EDGE[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                               {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};

// ---------------------------------------------------------------------------
// element kernel: one thread per tetrahedron
// ---------------------------------------------------------------------------
// Geometry of one tetrahedron (v: 4 vertices x 3): D = 6V, centroid g, and
// per local edge k the signed length ls_k = l_k si_k (orientation from the
// coordinate order x, then y, then z: tetNedelec_Direction,
// tetrahedral.c:225-263), the opposite-edge vector e_k = v_d - v_c
// (tetNedelec_mk, tetrahedral.c:267-281) and F_k = v_c x v_d + e_k x g, so
// that the edge function is N_k(r) = (ls_k / D)(v_c x v_d + e_k x r)
// (tetNedelec_km_mk / _Fxyz, tetrahedral.c:305-340, without the ls factor).
// Returns false for a degenerate or inverted element.
__device__ __forceinline__ bool tet_geom(const double v[12], double &D, double g[3], double ls[6], double e[6][3],
                                         double F[6][3]) {
  double r1[3], r2[3], r3[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    r1[c] = v[3 + c] - v[c];
    r2[c] = v[6 + c] - v[c];
    r3[c] = v[9 + c] - v[c];
  }
  // first row of the adjugate: D = r1 . (r2 x r3)
  const double G0 = r2[1] * r3[2] - r2[2] * r3[1];
  const double G1 = r2[2] * r3[0] - r2[0] * r3[2];
  const double G2 = r2[0] * r3[1] - r2[1] * r3[0];
  D = r1[0] * G0 + r1[1] * G1 + r1[2] * G2;
#pragma unroll
  for (int c = 0; c < 3; c++) g[c] = (v[c] + v[3 + c] + v[6 + c] + v[9 + c]) * 0.25;
  bool ok = D > 0.0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int ia = EDGE[k][0], ib = EDGE[k][1], ic = EDGE[k][2], id = EDGE[k][3];
    double dl[3], vc[3], vd[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      dl[c] = v[3 * ib + c] - v[3 * ia + c];
      vc[c] = v[3 * ic + c];
      vd[c] = v[3 * id + c];
      e[k][c] = vd[c] - vc[c];
    }
    double l = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    double si = dl[0] > 0.0 ? 1.0 : dl[0] < 0.0 ? -1.0
              : dl[1] > 0.0 ? 1.0 : dl[1] < 0.0 ? -1.0
              : dl[2] > 0.0 ? 1.0 : dl[2] < 0.0 ? -1.0 : 0.0;
    if (si == 0.0) ok = false;
    ls[k] = l * si;
    double w0 = vc[1] * vd[2] - vc[2] * vd[1];
    double w1 = vc[2] * vd[0] - vc[0] * vd[2];
    double w2 = vc[0] * vd[1] - vc[1] * vd[0];
    F[k][0] = w0 + e[k][1] * g[2] - e[k][2] * g[1];
    F[k][1] = w1 + e[k][2] * g[0] - e[k][0] * g[2];
    F[k][2] = w2 + e[k][0] * g[1] - e[k][1] * g[0];
  }
  return ok;
}

// G[i] = D grad(lambda_i): the columns of the adjugate of the edge matrix
__device__ __forceinline__ void tet_grads(const double v[12], double G[4][3]) {
  double r1[3], r2[3], r3[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    r1[c] = v[3 + c] - v[c];
    r2[c] = v[6 + c] - v[c];
    r3[c] = v[9 + c] - v[c];
  }
  G[1][0] = r2[1] * r3[2] - r2[2] * r3[1];
  G[1][1] = r2[2] * r3[0] - r2[0] * r3[2];
  G[1][2] = r2[0] * r3[1] - r2[1] * r3[0];
  G[2][0] = r3[1] * r1[2] - r3[2] * r1[1];
  G[2][1] = r3[2] * r1[0] - r3[0] * r1[2];
  G[2][2] = r3[0] * r1[1] - r3[1] * r1[0];
  G[3][0] = r1[1] * r2[2] - r1[2] * r2[1];
  G[3][1] = r1[2] * r2[0] - r1[0] * r2[2];
  G[3][2] = r1[0] * r2[1] - r1[1] * r2[0];
#pragma unroll
  for (int c = 0; c < 3; c++) G[0][c] = -(G[1][c] + G[2][c] + G[3][c]);
}

__device__ __forceinline__ void load_tet(const int *el_vtx, const double *crd, int el, double v[12]) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int g = el_vtx[4 * el + i];
#pragma unroll
    for (int c = 0; c < 3; c++) v[3 * i + c] = crd[3 * (long long)g + c];
  }
}

// centroid second moments C = (1/V) int rho rho^T dV = (1/20) sum (v_i - g)(v_i - g)^T
__device__ __forceinline__ void tet_moments(const double v[12], const double g[3], double C[3][3]) {
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) C[p][q] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double pp[3] = {v[3 * i] - g[0], v[3 * i + 1] - g[1], v[3 * i + 2] - g[2]};
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) C[p][q] += pp[p] * pp[q];
  }
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) C[p][q] *= 0.05;
}

// component c of the edge mass matrix (tetNedelec_sideside, tetrahedral.c:362-422):
// ls_i ls_j /(6D) (F_i,c F_j,c + E[(e_i x rho)_c (e_j x rho)_c]), dss = 1/(6D)
__device__ __forceinline__ double mass_comp(int i, int j, int c, const double ls[6], const double e[6][3],
                                            const double F[6][3], const double C[3][3], double dss) {
  const int b = (c + 1) % 3, cc = (c + 2) % 3;
  double q = e[i][b] * e[j][b] * C[cc][cc] - e[i][b] * e[j][cc] * C[cc][b] -
             e[i][cc] * e[j][b] * C[b][cc] + e[i][cc] * e[j][cc] * C[b][b];
  return dss * ls[i] * ls[j] * (F[i][c] * F[j][c] + q);
}

// uniaxial PML tensor of the complex stretch (sx, sy, sz): Lambda_c = s_b s_cc / s_c
__device__ __forceinline__ void pml_lambda(const double *pml, int el, double2 lam[3], double2 ilam[3]) {
  double2 s[3];
#pragma unroll
  for (int c = 0; c < 3; c++) s[c] = make_double2(pml[6 * (long long)el + 2 * c], pml[6 * (long long)el + 2 * c + 1]);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    lam[c] = cdiv_naive(cmul(s[(c + 1) % 3], s[(c + 2) % 3]), s[c]);
    ilam[c] = crecip_naive(lam[c]);
  }
}

__global__ __launch_bounds__(128) void k_element(AsmArgs a) {
  const int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= a.nel) return;
  double v[12];
  load_tet(a.el_vtx, a.crd, el, v);
  double D, g[3], ls[6], e[6][3], F[6][3];
  if (!tet_geom(v, D, g, ls, e, F)) {
    atomicOr(a.bad, 1);
    return;
  }
  double C[3][3];
  tet_moments(v, g, C);

  const double drr = 2.0 / (3.0 * D), dss = 1.0 / (6.0 * D);
  const double nu_omega = NU / a.omega;
  const double dd = -(a.eps[el] * a.omega);
  const double sig = a.sigma[el];
  double2 lam[3], ilam[3];
  const bool stretch = a.pml != nullptr;
  if (stretch) pml_lambda(a.pml, el, lam, ilam);
  double2 *out = a.elem_out + 27LL * el;
  int p = 0;
#pragma unroll
  for (int i = 0; i < 6; i++) {
#pragma unroll
    for (int j = i; j < 6; j++, p++) {
      double rr[3], mm[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        rr[c] = drr * ls[i] * ls[j] * e[i][c] * e[j][c];
        mm[c] = mass_comp(i, j, c, ls, e, F, C, dss);
      }
      double2 t;
      if (!stretch) {
        const double rsum = rr[0] + rr[1] + rr[2];
        const double ssum = mm[0] + mm[1] + mm[2];
        t.x = 0.0;
        t.x += nu_omega * rsum;
        t.x += dd * ssum;
        t.y = sig * ssum;
      } else {
        t = cz();
        const double2 epsc = make_double2(dd, sig);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          t = cadd(t, cscale(ilam[c], nu_omega * rr[c]));
          t = cadd(t, cscale(cmul(epsc, lam[c]), mm[c]));
        }
      }
      out[p] = t;
    }
  }
  if (a.dgv) {
    // delta-gap source (make_elem_em.c:320-396): vertex potentials i dV_j,
    // be_k = sigma omega sum_j (grad lambda_j, N_k) i dV_j with
    // (grad lambda_j, N_k) = ls_k F_k . G_j / (6D) (tetNedelec_gradside,
    // tetrahedral.c:645-661); the current source is not used in this mode
    double G[4][3];
    tet_grads(v, G);
    double dv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) dv[i] = a.dgv[a.el_vtx[4 * el + i]];
    const double dd6 = 1.0 / (6.0 * D), so = sig * a.omega;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double fx = ls[k] * F[k][0], fy = ls[k] * F[k][1], fz = ls[k] * F[k][2];
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) acc += dv[j] * (dd6 * (fx * G[j][0] + fy * G[j][1] + fz * G[j][2]));
      out[21 + k] = make_double2(0.0, acc * so);
    }
  } else if (a.src[el]) {
    // current source: exact integral of N_i . sum_j lambda_j J_j
    // (tetNedelec_JoA_Complex, lib/fem/tetrahedral.c:794-858)
    double G[4][3];
    tet_grads(v, G);
    double2 J[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      long long vb = 6LL * a.el_vtx[4 * el + i];
#pragma unroll
      for (int c = 0; c < 3; c++) J[i][c] = make_double2(a.joe[vb + 2 * c], a.joe[vb + 2 * c + 1]);
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int ia = EDGE[k][0], ib = EDGE[k][1];
      const double base = ls[k] / 120.0;
      double2 acc = cz();
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int jv = 0; jv < 4; jv++) {
          double pn = base * ((1.0 + (jv == ia)) * G[ib][c] - (1.0 + (jv == ib)) * G[ia][c]);
          acc = cadd(acc, cscale(J[jv][c], pn));
        }
      out[21 + k] = acc;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) out[21 + k] = cz();
  }
}

hipError_t launch_element(const AsmArgs &a, hipStream_t st) {
  if (a.nel <= 0) return hipSuccess;
  k_element<<<(a.nel + 127) / 128, 128, 0, st>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// element post-processing, one thread per tetrahedron
// (HDDM_post_em_E, main/fullwave/fullwave.c:640-763):
//   xe_k  = (nu / omega) x_c[edge k]                      (:721)
//   rotE  = (2/D)     sum_k (i xe_k) ls_k e_k             (:716-729)
//   E     = (2/D)     sum_k xe_k ls_k F_k(g)              (:733-748)
//   Je    = (sigma/D) sum_k xe_k ls_k F_k(g)              (:750-762)
// nu is 7.957747e5 for elements whose flag is a listed material and 0
// otherwise (:670-675); with N_k(r) = (ls_k/D)(v_c x v_d + e_k x r) these are
// rotE = i (nu/omega) curl E, E = 2 (nu/omega) E(g), Je = sigma (nu/omega) E(g),
// the reference's scalings kept as they are.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_post(PostArgs a) {
  const int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= a.nel) return;
  double v[12];
  load_tet(a.el_vtx, a.crd, el, v);
  double D, g[3], ls[6], e[6][3], F[6][3];
  (void)tet_geom(v, D, g, ls, e, F);
  const double s = (a.nu ? a.nu[el] : NU) / a.omega;
  double2 B[3] = {cz(), cz(), cz()}, E[3] = {cz(), cz(), cz()};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double2 x = cscale(a.X[a.el_dof[6 * (long long)el + k]], s);
    const double2 ix = make_double2(-x.y, x.x);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      B[c] = cadd(B[c], cscale(ix, ls[k] * e[k][c]));
      E[c] = cadd(E[c], cscale(x, ls[k] * F[k][c]));
    }
  }
  const double dd = 2.0 / D, ds = a.sigma[el] / D;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a.B[3 * (long long)el + c] = cscale(B[c], dd);
    a.E[3 * (long long)el + c] = cscale(E[c], dd);
    a.J[3 * (long long)el + c] = cscale(E[c], ds);
  }
}

hipError_t launch_post(const PostArgs &a, hipStream_t st) {
  if (a.nel <= 0) return hipSuccess;
  k_post<<<(a.nel + 127) / 128, 128, 0, st>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// adjoint sensitivity to the element permittivity, one thread per element.
// The element matrix depends on eps only through -eps*omega*M (M = SS, or
// sum_c Lambda_c MM_c in a stretched PML; make_elem_em.c:239-263 and the
// stretched form of k_element), so for J(E) with A E = f and A lambda =
// dJ/dE (A = A^T):  dJ/deps_e = -2 Re(lambda_e^T dA_e/deps E_e)
//                             =  2 omega Re(lambda_e^T M_e E_e).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_grad_eps(GradArgs a) {
  const int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= a.nel) return;
  double v[12];
  load_tet(a.el_vtx, a.crd, el, v);
  double D, g[3], ls[6], e[6][3], F[6][3];
  (void)tet_geom(v, D, g, ls, e, F);
  double C[3][3];
  tet_moments(v, g, C);
  const double dss = 1.0 / (6.0 * D);
  double2 lam[3], ilam[3];
  const bool stretch = a.pml != nullptr;
  if (stretch) pml_lambda(a.pml, el, lam, ilam);
  double2 xe[6], xl[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int d = a.el_dof[6 * (long long)el + k];
    xe[k] = a.xE[d];
    xl[k] = a.xL[d];
  }
  double2 acc = cz();
#pragma unroll
  for (int i = 0; i < 6; i++) {
    double2 row = cz();  // (M E)_i
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double2 m = cz();
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double mc = mass_comp(i, j, c, ls, e, F, C, dss);
        m = stretch ? cadd(m, cscale(lam[c], mc)) : make_double2(m.x + mc, 0.0);
      }
      row = cfma(row, m, xe[j]);
    }
    acc = cfma(acc, xl[i], row);
  }
  a.grad[el] = 2.0 * a.omega * acc.x;
}

hipError_t launch_grad_eps(const GradArgs &a, hipStream_t st) {
  if (a.nel <= 0) return hipSuccess;
  k_grad_eps<<<(a.nel + 127) / 128, 128, 0, st>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// assembly: out[i] = sum of in[src[k]] for k in ptr[i]..ptr[i+1]  (the order
// of the list is the reference's element order, so sums are accumulated in
// the order SetAe/SetFe add them)
// ---------------------------------------------------------------------------
__global__ void k_gather_sum(int n, const int *__restrict__ ptr, const int *__restrict__ src,
                             const double2 *__restrict__ in, double2 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc = cz();
  for (int k = ptr[i]; k < ptr[i + 1]; k++) acc = cadd(acc, in[src[k]]);
  out[i] = acc;
}
hipError_t launch_gather_sum(int n, const int *ptr, const int *src, const double2 *in,
                             double2 *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_gather_sum<<<(n + 255) / 256, 256, 0, st>>>(n, ptr, src, in, out);
  return hipGetLastError();
}

__global__ void k_copy_idx(int n, const int *__restrict__ slot, const double2 *__restrict__ in,
                           double2 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[slot[i]];
}
hipError_t launch_copy_idx(int n, const int *slot, const double2 *in, double2 *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_copy_idx<<<(n + 255) / 256, 256, 0, st>>>(n, slot, in, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dirichlet profile: first non-zero column of every row after SetBC +
// DeleteZero (aij_matrix_func.h:60-136, 342-391; advlas_skyline.c:12-25).
// Dirichlet rows become identity rows (first column = the row itself).
// ---------------------------------------------------------------------------
__global__ void k_profile(int ndof, const int *__restrict__ dof_sub, const SubMeta *__restrict__ meta,
                          const int *__restrict__ rowptr, const int *__restrict__ col,
                          const double2 *__restrict__ val, const int8_t *__restrict__ bcflag,
                          int *__restrict__ idx1) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndof) return;
  const int s = dof_sub[d];
  const int dof0 = meta[s].dof0;
  const int i = d - dof0;
  int first = i;
  if (bcflag[d] == 0) {
    for (int k = rowptr[d]; k < rowptr[d + 1]; k++) {
      const int j = col[k];
      if (bcflag[dof0 + j] == 0 && !ciszero(val[k])) {
        first = j;
        break;
      }
    }
  }
  idx1[d] = first;
}
hipError_t launch_profile(int ndof, const int *dof_sub, const SubMeta *meta, const int *rowptr,
                          const int *col, const double2 *val, const int8_t *bcflag, int *idx1,
                          hipStream_t st) {
  if (ndof <= 0) return hipSuccess;
  k_profile<<<(ndof + 255) / 256, 256, 0, st>>>(ndof, dof_sub, meta, rowptr, col, val, bcflag, idx1);
  return hipGetLastError();
}

// skyline values of the Dirichlet matrix (advlas_copy_nzmat_to_skymat, :49-60)
__global__ void k_skyfill(int nnz_tot, const int *__restrict__ ent_row, const int *__restrict__ dof_sub,
                          const SubMeta *__restrict__ meta, const int *__restrict__ col,
                          const double2 *__restrict__ val, const int8_t *__restrict__ bcflag,
                          const int *__restrict__ idx1, const int *__restrict__ rowbeg,
                          double2 *__restrict__ fac) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz_tot) return;
  const int d = ent_row[k];
  if (bcflag[d] != 0) return;
  const int s = dof_sub[d];
  const SubMeta m = meta[s];
  const int j = col[k];
  if (bcflag[m.dof0 + j] != 0) return;
  const double2 v = val[k];
  if (ciszero(v)) return;
  fac[m.fac0 + rowbeg[d] + (j - idx1[d])] = v;
}
hipError_t launch_skyfill(int nnz_tot, const int *ent_row, const int *dof_sub, const SubMeta *meta,
                          const int *col, const double2 *val, const int8_t *bcflag,
                          const int *idx1, const int *rowbeg, double2 *fac, hipStream_t st) {
  if (nnz_tot <= 0) return hipSuccess;
  k_skyfill<<<(nnz_tot + 255) / 256, 256, 0, st>>>(nnz_tot, ent_row, dof_sub, meta, col, val, bcflag,
                                                  idx1, rowbeg, fac);
  return hipGetLastError();
}

// DOFs [d0, d0 + ndof)
__global__ void k_skydiag_bc(int d0, int ndof, const int *__restrict__ dof_sub, const SubMeta *__restrict__ meta,
                             const int8_t *__restrict__ bcflag, const int *__restrict__ rowbeg,
                             const int *__restrict__ idx1, double2 *__restrict__ fac) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndof) return;
  const int d = d0 + t;
  if (bcflag[d] == 0) return;
  const SubMeta m = meta[dof_sub[d]];
  const int i = d - m.dof0;
  fac[m.fac0 + rowbeg[d] + (i - idx1[d])] = make_double2(1.0, 0.0);
}
hipError_t launch_skydiag_bc(int d0, int ndof, const int *dof_sub, const SubMeta *meta, const int8_t *bcflag,
                             const int *rowbeg, const int *idx1, double2 *fac, hipStream_t st) {
  if (ndof <= 0) return hipSuccess;
  k_skydiag_bc<<<(ndof + 255) / 256, 256, 0, st>>>(d0, ndof, dof_sub, meta, bcflag, rowbeg, idx1, fac);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// skyline LDL^T (advlas_skyline_ldl_decomposite, advlas_skyline.c:79-128):
// one wavefront per subdomain, row by row.  Row i is built in LDS:
//   u_ij = a_ij - sum_{k=max(f_i,f_j)}^{j-1} u_ik l_jk      (j ascending)
//   d_i  = a_ii - sum_k u_ik^2 / d_k,  l_ik = u_ik / d_k
// and written back; finished rows are re-read through L2 (agent-scope
// acquire after every row), the pivots 1/d_k stay in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_factor(FactorArgs a) {
  const int w = a.order ? a.order[a.base + blockIdx.x] : a.base + blockIdx.x;
  const SubMeta m = a.meta[w];
  extern __shared__ double2 lds[];
  double2 *u = lds;              // [wmax+1]
  double2 *dg = lds + a.wmax + 1;  // [nmax]
  const int lane = threadIdx.x;
  const int *idx1 = a.idx1 + m.dof0;
  const int *rowbeg = a.rowbeg + m.dof0;
  double2 *S = a.fac + m.fac0;
  for (int i = 0; i < m.n; i++) {
    const int fi = idx1[i];
    const int rb = rowbeg[i];
    const int wi = i - fi;
    for (int t = lane; t <= wi; t += 64) u[t] = S[rb + t];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    for (int j = fi; j < i; j++) {
      const int fj = idx1[j];
      const int rbj = rowbeg[j];
      const int k0 = fi > fj ? fi : fj;
      double2 acc = cz();
      for (int k = k0 + lane; k < j; k += 64) acc = cfma(acc, u[k - fi], S[rbj + (k - fj)]);
      acc = wave_csum(acc);
      if (lane == 0) u[j - fi] = csub(u[j - fi], acc);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    double2 r = cz();
    for (int k = fi + lane; k < i; k += 64) {
      const double2 uk = u[k - fi];
      const double2 dk = dg[k];
      r = cadd(r, cmul(cmul(uk, uk), dk));
      u[k - fi] = cmul(uk, dk);
    }
    r = wave_csum(r);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) {
      const double2 piv = crecip_smith(csub(u[wi], r));
      dg[i] = piv;
      u[wi] = piv;
      a.dinv[m.dof0 + i] = piv;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    for (int t = lane; t <= wi; t += 64) S[rb + t] = u[t];
    // make the finished row visible to this wave's later loads
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
  }
}
hipError_t launch_factor(const FactorArgs &a, int nwork, hipStream_t st) {
  if (nwork <= 0) return hipSuccess;
  size_t lds = sizeof(double2) * (size_t)(a.wmax + 1 + a.nmax);
  k_factor<<<nwork, 64, lds, st>>>(a);
  return hipGetLastError();
}


// diagonal of A at every slot, zero -> 1 (hddm_pc_CS_DIAG_make, hddm_pc.h:52-86)
__global__ void k_pcslot(int nslots, const int *__restrict__ diag_ent, const int *__restrict__ unused,
                         const double2 *__restrict__ val, double2 *__restrict__ pcslot) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nslots) return;
  const int e = diag_ent[k];
  double2 d = e >= 0 ? val[e] : cz();
  if (ciszero(d)) d = make_double2(1.0, 0.0);
  pcslot[k] = d;
}
hipError_t launch_pcslot(int nslots, const int *diag_ent, const int *unused, const double2 *val,
                         double2 *pcslot, hipStream_t st) {
  if (nslots <= 0) return hipSuccess;
  k_pcslot<<<(nslots + 255) / 256, 256, 0, st>>>(nslots, diag_ent, unused, val, pcslot);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// interface vectors
// ---------------------------------------------------------------------------
// gather of per-subdomain slot values into interface unknowns, in the
// reference's order: within a part, subdomains in order (COM_res_sp,
// hddm_solver_sp.h:41-54); parts summed owner-first, ascending
// (hddm_solver_s.h:404-424)
__device__ __forceinline__ double2 gather_slots(const int *__restrict__ ptr, const int *__restrict__ slot,
                                                const int4 *__restrict__ slot4, const double2 *__restrict__ y,
                                                int g) {
  double2 acc = cz(), part = cz();
  const int4 s4 = slot4[g];
  if (!(s4.x == SLOT_NONE && s4.y == SLOT_MORE)) {
    // the usual case (an interface edge in 2-4 subdomains): the slot ids
    // come in one load, the values are requested together, then summed in
    // list order
    int sl[4] = {s4.x, s4.y, s4.z, s4.w};
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) cnt += sl[j] != SLOT_NONE;
    double2 v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = j < cnt ? y[sl[j] < 0 ? ~sl[j] : sl[j]] : cz();
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j < cnt) {
        if (sl[j] < 0) {
          if (j != 0) acc = cadd(acc, part);
          part = v[j];
        } else {
          part = cadd(part, v[j]);
        }
      }
    }
    return cnt > 0 ? cadd(acc, part) : acc;
  }
  const int c0 = ptr[g], c1 = ptr[g + 1];
  for (int c = c0; c < c1; c++) {
    int sl = slot[c];
    if (sl < 0) {
      sl = ~sl;
      if (c != c0) acc = cadd(acc, part);
      part = y[sl];
    } else {
      part = cadd(part, y[sl]);
    }
  }
  return c1 > c0 ? cadd(acc, part) : acc;
}

constexpr int VT = 256;

__device__ __forceinline__ double2 block_csum(double2 v, double2 *sh) {
  v = wave_csum(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double2 t = cz();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < VT / 64; k++) t = cadd(t, sh[k]);
    sh[VT / 64] = t;
  }
  __syncthreads();
  return sh[VT / 64];
}
__device__ __forceinline__ double block_max(double v, double *sh) {
  v = wave_max(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
#pragma unroll
    for (int k = 1; k < VT / 64; k++) t = fmax(t, sh[k]);
    sh[VT / 64] = t;
  }
  __syncthreads();
  return sh[VT / 64];
}
// identical fixed-order reduction of G partials in every block
// (eight loads per thread in flight, added in the same order)
__device__ __forceinline__ double2 reduce_parts(const double2 *__restrict__ p, int G, double2 *sh) {
  double2 t = cz();
  for (int k = threadIdx.x; k < G; k += 8 * VT) {
    double2 v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = k + j * VT < G ? p[k + j * VT] : cz();
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (k + j * VT < G) t = cadd(t, v[j]);
  }
  return block_csum(t, sh);
}
__device__ __forceinline__ double reduce_parts_max(const double2 *__restrict__ p, int G, double *sh) {
  double t = 0.0;
  for (int k = threadIdx.x; k < G; k += 8 * VT) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = k + j * VT < G ? p[k + j * VT].x : 0.0;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (k + j * VT < G) t = fmax(t, v[j]);
  }
  return block_max(t, sh);
}

// the vectors of system sys0 + blockIdx.y (index arrays are shared)
__device__ __forceinline__ void to_system(VecArgs &a) {
  const long long s = a.sys0 + (long long)blockIdx.y;
  if (s == 0) return;
  a.yslot += s * a.ss.slot;
  if (a.pcslot) a.pcslot += s * a.ss.pcs;
  a.P += s * a.ss.g;
  a.Q += s * a.ss.g;
  a.R += s * a.ss.g;
  a.Z += s * a.ss.g;
  a.U += s * a.ss.g;
  a.PC += s * a.ss.g;
  a.part += s * 3LL * a.G;
  if (a.pqpart) a.pqpart += s * (long long)a.npq;
  a.sc += s;
  a.hist += s * a.ss.hist;
}
__host__ __forceinline__ dim3 vec_grid(const VecArgs &a, int gx) { return dim3(gx, a.nsys > 0 ? a.nsys : 1); }

// first step: Q = gather(y), PC = 1/gather(diag) (hddm_pc.h:108-137),
// R = Q_owned, Z = R PC, P = Z, U = 0 (hddm_solver_s.h:158-256)
__global__ __launch_bounds__(VT) void k_first_a(VecArgs a) {
  to_system(a);
  __shared__ double2 shc[VT / 64 + 1];
  __shared__ double shm[VT / 64 + 1];
  double2 rz = cz();
  double mx = 0.0;
  for (int g = blockIdx.x * VT + threadIdx.x; g < a.ng; g += a.G * VT) {
    const double2 q = gather_slots(a.gq_ptr, a.gq_slot, a.gq_slot4, a.yslot, g);
    const double2 ps = gather_slots(a.gq_ptr, a.gq_slot, a.gq_slot4, a.pcslot, g);
    const double2 pc = ciszero(ps) ? make_double2(1.0, 0.0) : cdiv_naive(make_double2(1.0, 0.0), ps);
    const double2 z = cmul(q, pc);
    a.Q[g] = q;
    a.PC[g] = pc;
    a.R[g] = q;
    a.Z[g] = z;
    a.P[g] = z;
    a.U[g] = cz();
    rz = cadd(rz, cmul(q, z));
    mx = fmax(mx, cabsv(q));
  }
  rz = block_csum(rz, shc);
  mx = block_max(mx, shm);
  if (threadIdx.x == 0) {
    a.part[a.G + blockIdx.x] = rz;
    a.part[2 * a.G + blockIdx.x] = make_double2(mx, 0.0);
  }
}
__global__ __launch_bounds__(VT) void k_first_b(VecArgs a) {
  to_system(a);
  __shared__ double2 shc[VT / 64 + 1];
  __shared__ double shm[VT / 64 + 1];
  const double2 rz = reduce_parts(a.part + a.G, a.G, shc);
  const double r0 = reduce_parts_max(a.part + 2 * a.G, a.G, shm);
  if (threadIdx.x == 0) {
    Scal *s = a.sc;
    s->done = (r0 < NEARLY_ZERO) ? 1 : 0;
    s->r0inv = 1.0 / r0;
    s->rz = rz;
    s->beta = cz();  // p_1 = z_0 + 0 p, with P = Z already
    s->loop = 1;
    s->iters = 0;
    s->conv = 0;
    s->resi = 1.0;
    s->pending = 0;
    s->st_rz = rz;
    s->st_beta = cz();
    s->st_done = s->done;
    s->st_fresh = 0;
    a.hist[0] = 1.0;
  }
}
hipError_t launch_first(const VecArgs &a, hipStream_t st) {
  k_first_a<<<vec_grid(a, a.G), VT, 0, st>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_first_b<<<vec_grid(a, 1), VT, 0, st>>>(a);
  return hipGetLastError();
}

// Interface iteration (hddm_solver_s.h:260-400) after the Schur step (which
// leaves y per slot and every subdomain's part of (p,q)): one pass over the
// unknowns and the scalar status step.  Every reduction sums per-block or
// per-subdomain partials in one fixed order, so results do not depend on
// scheduling; no kernel writes a scalar that another block of the same kernel
// reads.
//   k_iter_ab: alpha from the (p,q) parts (every block, same order,
//              :286-295), p_k = z_k + beta p_{k-1} (written back to P),
//              q = gather(y), u += alpha p, r -= alpha q, z = PC r, partials
//              of (r,z) and max|r|
//   k_iter_c:  one block: beta (:340-350), the stop test (:311-318,
//              :370-376) and the status machine (hddm_solver.h:125-143)
// The Schur step of the next iteration forms p = z + beta p itself (zbp,
// same fused operations as k_iter_ab); the final P (U when converged,
// hddm_solver_s.h:370-376) is set by k_finalize_p before the last solve.
// Inside a graph of explicit iterations k_iter_c runs only after the last
// one: the explicit Schur matvec takes the status step from k_iter_ab's
// partials itself (every workgroup, the same reduction order) and stages it,
// and the next k_iter_ab's block 0 commits it (Scal::pending, st_*).
__global__ __launch_bounds__(VT) void k_iter_c(VecArgs a) {
  to_system(a);
  Scal *s = a.sc;
  __shared__ double2 shc[VT / 64 + 1];
  __shared__ double shm[VT / 64 + 1];
  // every scalar is requested with the partials (one round trip); the
  // reductions do not depend on the status, which is tested after them (a
  // uniform branch)
  const int done = s->done, pending = s->pending, loop = s->loop, iters = s->iters, max_loop = s->max_loop;
  const double r0inv = s->r0inv, eps = s->eps;
  const double2 rzo = s->rz;
  const double2 rzs = reduce_parts(a.part + a.G, a.G, shc);
  const double mxs = reduce_parts_max(a.part + 2 * a.G, a.G, shm);
  if (done || !pending) return;
  if (threadIdx.x == 0) {
    const StatusStep r = status_step(rzs, mxs, rzo, r0inv, eps, loop, max_loop);
    a.hist[loop] = r.resi;
    s->resi = r.resi;
    s->conv = r.conv;
    s->rz = rzs;
    s->beta = r.beta;
    s->iters = iters + 1;
    s->done = r.done;
    s->loop = loop + 1;
    s->pending = 0;
    s->st_rz = rzs;
    s->st_beta = r.beta;
    s->st_done = r.done;
    s->st_fresh = 0;
  }
}
// The vector pass of an interface iteration.  The Schur step leaves every
// subdomain's part of (p,q) (pqpart: the sum over its slots of p y, which
// regroups sum_g p_g q_g by subdomain), so alpha is known when the pass
// starts.  The status comes from the staged fields only
// (st_rz, st_beta, st_done: equal to the committed ones when nothing is
// staged), which no block of this kernel writes; block 0 commits a step the
// matvec staged.
__global__ __launch_bounds__(VT) void k_iter_ab(VecArgs a) {
  to_system(a);
  const int g0 = blockIdx.x * VT + threadIdx.x;
  double2 q0 = cz(), z0 = cz(), p0 = cz(), u0 = cz(), r0 = cz(), pc0 = cz();
  if (g0 < a.ng) {
    q0 = gather_slots(a.gq_ptr, a.gq_slot, a.gq_slot4, a.yslot, g0);
    z0 = a.Z[g0];
    p0 = a.P[g0];
    u0 = a.U[g0];
    r0 = a.R[g0];
    pc0 = a.PC[g0];
  }
  Scal *s = a.sc;
  const int done = s->st_done;
  const double2 beta = s->st_beta, rz = s->st_rz;
  if (blockIdx.x == 0 && threadIdx.x == 0 && s->st_fresh) {
    const int loop = s->loop;
    const double resi = s->st_resi;
    a.hist[loop] = resi;
    s->resi = resi;
    s->conv = s->st_conv;
    s->rz = s->st_rz;
    s->beta = s->st_beta;
    s->iters = s->iters + 1;
    s->done = s->st_done;
    s->loop = loop + 1;
    s->st_fresh = 0;
    s->pending = 0;
  }
  __shared__ double2 shc[VT / 64 + 1];
  __shared__ double shm[VT / 64 + 1];
  const double2 pq = reduce_parts(a.pqpart, a.npq, shc);
  if (done) return;
  // this iteration's status step is pending until k_iter_c or the next
  // Schur matvec takes it (no other block of this kernel reads the flag)
  if (blockIdx.x == 0 && threadIdx.x == 0) s->pending = 1;
  double2 malpha = cz();
  if (!(cabsv(rz) < NEARLY_ZERO && cabsv(pq) < NEARLY_ZERO)) malpha = csub(cz(), cdiv_naive(rz, pq));
  double2 rzn = cz();
  double mx = 0.0;
  if (g0 < a.ng) {
    const double2 p = zbp(z0, beta, p0);
    a.P[g0] = p;
    const double2 u = cadd(u0, cmul(malpha, p));
    const double2 r = cadd(r0, cmul(malpha, q0));
    const double2 z = cmul(r, pc0);
    a.U[g0] = u;
    a.R[g0] = r;
    a.Z[g0] = z;
    mx = cabsv(r);
    rzn = cmul(r, z);
  }
  for (int g = g0 + a.G * VT; g < a.ng; g += a.G * VT) {
    const double2 q = gather_slots(a.gq_ptr, a.gq_slot, a.gq_slot4, a.yslot, g);
    const double2 p = zbp(a.Z[g], beta, a.P[g]);
    a.P[g] = p;
    const double2 u = cadd(a.U[g], cmul(malpha, p));
    const double2 r = cadd(a.R[g], cmul(malpha, q));
    const double2 z = cmul(r, a.PC[g]);
    a.U[g] = u;
    a.R[g] = r;
    a.Z[g] = z;
    mx = fmax(mx, cabsv(r));
    rzn = cadd(rzn, cmul(r, z));
  }
  rzn = block_csum(rzn, shc);
  mx = block_max(mx, shm);
  if (threadIdx.x == 0) {
    a.part[a.G + blockIdx.x] = rzn;
    a.part[2 * a.G + blockIdx.x] = make_double2(mx, 0.0);
  }
}
hipError_t launch_iter_ab(const VecArgs &a, hipStream_t st, bool status_step) {
  k_iter_ab<<<vec_grid(a, a.G), VT, 0, st>>>(a);
  if (status_step) k_iter_c<<<vec_grid(a, 1), VT, 0, st>>>(a);
  return hipGetLastError();
}


// P for the last solve: U when the stop test fired, else the next search
// direction z + beta p (max_loop reached, hddm_solver.h:125-143)
__global__ __launch_bounds__(VT) void k_finalize_p(VecArgs a) {
  to_system(a);
  const Scal *s = a.sc;
  const bool conv = s->conv != 0;
  const double2 beta = s->beta;
  for (int g = blockIdx.x * VT + threadIdx.x; g < a.ng; g += a.G * VT)
    a.P[g] = conv ? a.U[g] : zbp(a.Z[g], beta, a.P[g]);
}
hipError_t launch_finalize_p(const VecArgs &a, hipStream_t st) {
  k_finalize_p<<<vec_grid(a, a.G), VT, 0, st>>>(a);
  return hipGetLastError();
}

__global__ __launch_bounds__(VT) void k_gather_q(VecArgs a) {
  to_system(a);
  for (int g = blockIdx.x * VT + threadIdx.x; g < a.ng; g += a.G * VT)
    a.Q[g] = gather_slots(a.gq_ptr, a.gq_slot, a.gq_slot4, a.yslot, g);
}
hipError_t launch_gather_q(const VecArgs &a, hipStream_t st) {
  k_gather_q<<<vec_grid(a, a.G), VT, 0, st>>>(a);
  return hipGetLastError();
}

__global__ void k_fill(double2 *p, long long n, double2 v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
hipError_t launch_fill(double2 *p, long long n, double2 v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  k_fill<<<(int)nb, 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

__global__ void k_scal_init(Scal *s, double eps, int max_loop) {
  s += blockIdx.x;
  s->eps = eps;
  s->max_loop = max_loop;
  s->done = 0;
  s->conv = 0;
  s->loop = 0;
  s->iters = 0;
  s->beta = cz();
  s->pending = 0;
  s->st_rz = cz();
  s->st_beta = cz();
  s->st_done = 0;
  s->st_fresh = 0;
}
hipError_t launch_scal_init(Scal *sc, int nsys, double eps, int max_loop, hipStream_t st) {
  k_scal_init<<<nsys, 1, 0, st>>>(sc, eps, max_loop);
  return hipGetLastError();
}

}  // namespace fw

namespace fw {

// ---------------------------------------------------------------------------
// IC(0) factor of every subdomain's Dirichlet matrix (aij_pc_func.h:36-108),
// one wavefront per subdomain, row by row:
//   L_ik = a_ik - sum_{c<k in row i} d_c L_ic L_kc   (entries of row i in order)
//   L_ii = af a_ii - sum_c d_c L_ic^2,   d_i = 1 / L_ii
// L_kc is looked up in the finished row k (binary search; zero when (k,c) is
// outside the pattern, as getAIJ returns).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 lookup_row(const int *__restrict__ col, const double2 *L, int b, int e, int c) {
  int lo = b, hi = e;  // [lo, hi)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int cm = col[mid];
    if (cm == c) return L[mid];
    if (cm < c) lo = mid + 1; else hi = mid;
  }
  return cz();
}

__global__ __launch_bounds__(64) void k_icc_make(IccArgs a) {
  const SubMeta m = a.meta[blockIdx.x];
  extern __shared__ double2 lds[];
  double2 *dgl = lds;           // [nmax]
  double2 *u = lds + a.nmax;    // current row
  const int lane = threadIdx.x;
  for (int i = 0; i < m.n; i++) {
    const int d = m.dof0 + i;
    const int ii = a.dptr[d], jj = a.dptr[d + 1] - 1;
    for (int t = lane; t <= jj - ii; t += 64) u[t] = a.dval[ii + t];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    for (int row = ii; row < jj; row++) {
      const int k = a.dcol[row];
      const int kb = a.dptr[m.dof0 + k], ke = a.dptr[m.dof0 + k + 1];
      double2 acc = cz();
      for (int t = ii + lane; t < row; t += 64) {
        const int c = a.dcol[t];
        const double2 lkc = lookup_row(a.dcol, a.L, kb, ke, c);
        acc = cadd(acc, cmul(dgl[c], cmul(u[t - ii], lkc)));
      }
      acc = wave_csum(acc);
      if (lane == 0) u[row - ii] = csub(a.dval[row], acc);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    double2 acc = cz();
    for (int t = ii + lane; t < jj; t += 64) {
      const double2 q = u[t - ii];
      acc = cadd(acc, cmul(dgl[a.dcol[t]], cmul(q, q)));
    }
    acc = wave_csum(acc);
    if (lane == 0) {
      const double2 ljj = csub(cscale(a.dval[jj], a.af), acc);
      u[jj - ii] = ljj;
      const double2 di = cdiv_naive(make_double2(1.0, 0.0), ljj);
      dgl[i] = di;
      a.dg[d] = di;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    for (int t = lane; t <= jj - ii; t += 64) a.L[ii + t] = u[t];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
  }
}
hipError_t launch_icc_make(const IccArgs &a, int nsub, hipStream_t st) {
  if (nsub <= 0) return hipSuccess;
  // row buffer: the longest Dirichlet row is at most nmax entries
  k_icc_make<<<nsub, 64, sizeof(double2) * 2 * (size_t)a.nmax, st>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// COCG with scaled search directions and ICC (aij_solver_cg.h:113-355):
//   r = b - A x, q = v/(v,r) with v = M^-1 r, then
//   alpha = 1/(q,Aq), x += alpha q, r -= alpha Aq, stop on ||r||_2/||b||_2,
//   q += v/(v,r).   Bilinear (unconjugated) products in the recurrences.
// One 256-thread workgroup per subdomain, vectors in LDS; the triangular
// ICC solves run in wavefront 0.
// ---------------------------------------------------------------------------
constexpr int CT = 256;

__device__ __forceinline__ double2 bsum(double2 v, double2 *sh) {
  v = wave_csum(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double2 t = cz();
#pragma unroll
  for (int k = 0; k < CT / 64; k++) t = cadd(t, sh[k]);
  __syncthreads();
  return t;
}

__device__ void icc_apply(const IccArgs &a, const SubMeta &m, const double2 *r, double2 *v) {
  // forward: v_i = (r_i - sum_j v_j L_ij) / L_ii        (aij_pc_func.h:135-145)
  // backward: v_c -= v_i L_ic d_c, i descending          (aij_pc_func.h:148-156)
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    for (int i = 0; i < m.n; i++) {
      const int d = m.dof0 + i;
      const int ii = a.dptr[d], jj = a.dptr[d + 1] - 1;
      double2 acc = cz();
      for (int t = ii + lane; t < jj; t += 64) acc = cadd(acc, cmul(v[a.dcol[t]], a.L[t]));
      acc = wave_csum(acc);
      if (lane == 0) v[i] = cdiv_naive(csub(r[i], acc), a.L[jj]);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    for (int i = m.n - 1; i >= 0; i--) {
      const int d = m.dof0 + i;
      const int ii = a.dptr[d], jj = a.dptr[d + 1] - 1;
      const double2 vi = v[i];
      for (int t = ii + lane; t < jj; t += 64) {
        const int c = a.dcol[t];
        v[c] = csub(v[c], cmul(vi, cmul(a.L[t], a.dg[m.dof0 + c])));
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void full_mxv(const IccArgs &a, const SubMeta &m, const double2 *in, double2 *out) {
  for (int i = threadIdx.x; i < m.n; i += CT) {
    const int d = m.dof0 + i;
    double2 acc = cz();
    for (int t = a.fptr[d]; t < a.fptr[d + 1]; t++) acc = cfma(acc, a.fval[t], in[a.fcol[t]]);
    out[i] = acc;
  }
  __syncthreads();
}

__global__ __launch_bounds__(CT) void k_cocg(IccArgs a) {
  const SubMeta m = a.meta[blockIdx.x];
  extern __shared__ double2 lds[];
  double2 *x = lds, *r = x + a.nmax, *v = r + a.nmax, *q = v + a.nmax, *aq = q + a.nmax;
  __shared__ double2 sh[CT / 64];
  const int n = m.n;
  int status = -1;
  for (int i = threadIdx.x; i < n; i += CT) x[i] = a.x[m.dof0 + i];
  __syncthreads();
  full_mxv(a, m, x, aq);
  double2 bb = cz(), rr = cz();
  for (int i = threadIdx.x; i < n; i += CT) {
    const double2 bi = a.b[m.dof0 + i];
    r[i] = csub(bi, aq[i]);
    bb = cadd(bb, make_double2(cabs2(bi), 0.0));
    rr = cadd(rr, make_double2(cabs2(r[i]), 0.0));
  }
  bb = bsum(bb, sh);
  rr = bsum(rr, sh);
  double bnorm = sqrt(bb.x);
  if (bnorm < NEARLY_ZERO) {
    status = 0;
  } else {
    bnorm = 1.0 / bnorm;
    double res = sqrt(rr.x) * bnorm;
    if (res <= a.eps) {
      status = 0;
    } else {
      icc_apply(a, m, r, v);
      double2 vr = cz();
      for (int i = threadIdx.x; i < n; i += CT) vr = cfma(vr, v[i], r[i]);
      double2 rv = crecip_naive(bsum(vr, sh));
      for (int i = threadIdx.x; i < n; i += CT) q[i] = cmul(v[i], rv);
      __syncthreads();
      const int max_ite = 10 * n;
      status = -3;  // SOLVER_STATUS_REACHED_MAX_ITE
      for (int it = 1; it < max_ite; it++) {
        full_mxv(a, m, q, aq);
        double2 qa = cz();
        for (int i = threadIdx.x; i < n; i += CT) qa = cfma(qa, q[i], aq[i]);
        double2 alpha = bsum(qa, sh);
        if (ciszero(alpha)) { status = -5; break; }
        alpha = cdiv_naive(make_double2(1.0, 0.0), alpha);
        const double2 malpha = csub(cz(), alpha);
        double2 r2 = cz();
        for (int i = threadIdx.x; i < n; i += CT) {
          x[i] = cadd(x[i], cmul(alpha, q[i]));
          r[i] = cadd(r[i], cmul(malpha, aq[i]));
          r2 = cadd(r2, make_double2(cabs2(r[i]), 0.0));
        }
        r2 = bsum(r2, sh);
        res = sqrt(r2.x) * bnorm;
        if (res <= a.eps) { status = it; break; }
        if (res >= a.div) { status = -2; break; }
        icc_apply(a, m, r, v);
        vr = cz();
        for (int i = threadIdx.x; i < n; i += CT) vr = cfma(vr, v[i], r[i]);
        rv = crecip_naive(bsum(vr, sh));
        for (int i = threadIdx.x; i < n; i += CT) q[i] = cadd(q[i], cmul(v[i], rv));
        __syncthreads();
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += CT) a.x[m.dof0 + i] = x[i];
  if (threadIdx.x == 0) a.status[blockIdx.x] = status;
}
hipError_t launch_cocg(const IccArgs &a, int nsub, hipStream_t st) {
  if (nsub <= 0) return hipSuccess;
  k_cocg<<<nsub, CT, sizeof(double2) * 5 * (size_t)a.nmax, st>>>(a);
  return hipGetLastError();
}

}  // namespace fw
