// fw_factor.hip - blocked skyline LDL^T factorisation on the FP64 matrix
// cores (gfx950).
//
// The reference factors each subdomain's Dirichlet matrix row by row
// (advlas_skyline_ldl_decomposite, lib/hddm/advlas_skyline.c:79-128; row
// Crout, no pivoting, no conjugation):
//   u_ij = a_ij - sum_{k=max(f_i,f_j)}^{j-1} u_ik l_jk   (f_i <= j < i)
//   d_i  = a_ii - sum_k u_ik l_ik,   l_ik = u_ik / d_k
// and stores l_ik in row i's profile entries and 1/d_i on the diagonal.
//
// Here one wavefront factors one subdomain in uniform blocks of 16 rows,
// left-looking.  For block row R and each column block C < R reached by its
// profile:
//   X = A[R,C] - sum_K U[R,K] L[C,K]^T       (k-steps of 4 columns)
//   U[R,C] = X M_C^T,  M_C = L_CC^-1          (the unit lower diagonal block)
// and for the diagonal block X = A[R,R] - sum_K U[R,K] (U[R,K] D_K^-1)^T,
// factored in LDS row by row of pivots (right-looking inside the block).
// Until block row R is finished its profile entries hold the unscaled u;
// then they are scaled by 1/d of their columns, as the reference scales a
// finished row.  Entries outside the profile come out exactly zero, so only
// the profile is read and written.  All products run on
// v_mfma_f64_16x16x4f64 (A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// D[i][j] in lane j + 16 (i mod 4), register i / 4); a complex product is
// four real ones into separate accumulators.
#include "fw_kernels.h"
#include "fw_device.h"

namespace fw {

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma64(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(64) void k_factor16(FactorArgs a) {
  const int w = a.order ? a.order[a.base + blockIdx.x] : a.base + blockIdx.x;
  const SubMeta m = a.meta[w];
  const int n = m.n;
  const int *idx1 = a.idx1 + m.dof0, *rowbeg = a.rowbeg + m.dof0;
  double2 *S = a.fac + m.fac0;
  double2 *dinv = a.dinv + m.dof0;
  double2 *Mall = a.m16 + a.mf_off[w] * 256;
  __shared__ double2 xs[16][17];  // a 16x16 block (transposes, the diagonal block)
  __shared__ double2 pv[16];      // pivots 1/d of the block row
  const int l = threadIdx.x, li = l & 15, lk = l >> 4;
  const mfma_d4 z4 = {0.0, 0.0, 0.0, 0.0};

  for (int r0 = 0; r0 < n; r0 += 16) {
    const int R = r0 >> 4, nb = min(16, n - r0);
    // row li of the block (A operand rows, and the B rows of the diagonal)
    const bool rv = li < nb;
    const int rowA = r0 + (rv ? li : 0);
    const int fA = rv ? idx1[rowA] : r0, rbA = rowbeg[rowA];
    // rows 4v + lk of the block (D layout)
    int fD[4], rbD[4];
    bool vD[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = 4 * v + lk, row = r0 + (i < nb ? i : 0);
      vD[v] = i < nb;
      fD[v] = idx1[row];
      rbD[v] = rowbeg[row];
    }
    int cmin = fA;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) cmin = min(cmin, __shfl_xor(cmin, o, 64));
    const int cb0 = cmin >> 4;

    for (int C = cb0; C <= R; C++) {
      const int c0 = 16 * C;
      const bool diag = C == R;
      // X = A[R, C] (D layout: register v holds row 4v + lk, column c0 + li)
      mfma_d4 xr, xi;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = r0 + 4 * v + lk, j = c0 + li;
        const bool in = vD[v] && j >= fD[v] && j <= i;
        const double2 x = in ? S[rbD[v] + (j - fD[v])] : cz();
        xr[v] = x.x;
        xi[v] = x.y;
      }
      // rows of block C as the B operand (off-diagonal blocks)
      const int rowB = c0 + li;  // < n: only the last block can be partial
      const int fB = diag ? fA : idx1[rowB], rbB = diag ? rbA : rowbeg[rowB];
      int kmin = fB;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) kmin = min(kmin, __shfl_xor(kmin, o, 64));
      const int kb0 = max(cb0, kmin >> 4);
      mfma_d4 arr = z4, aii = z4, ari = z4, air = z4;
      for (int K = kb0; K < C; K++) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int k = 16 * K + 4 * s + lk;
          const double2 A = (rv && k >= fA) ? S[rbA + (k - fA)] : cz();  // U[R,K][li][k]
          double2 B;
          if (diag) {
            B = cmul(A, dinv[k]);  // L[R,K][li][k] = u l-scaled on the fly
          } else {
            B = k >= fB ? S[rbB + (k - fB)] : cz();  // L[C,K][li][k], finished
          }
          arr = mfma64(A.x, B.x, arr);
          aii = mfma64(A.y, B.y, aii);
          ari = mfma64(A.x, B.y, ari);
          air = mfma64(A.y, B.x, air);
        }
      }
      xr = xr - (arr - aii);
      xi = xi - (ari + air);
#pragma unroll
      for (int v = 0; v < 4; v++) xs[4 * v + lk][li] = make_double2(xr[v], xi[v]);
      __syncthreads();
      if (!diag) {
        // U[R,C] = X M_C^T: A[i][k] = X[i][4s+k], B[k][j] = M_C[j][4s+k]
        const double2 *Mc = Mall + (long long)C * 256 + li * 16 + lk;
        mfma_d4 orr = z4, oii = z4, ori = z4, oir = z4;
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const double2 A = xs[li][4 * s + lk];
          const double2 B = Mc[4 * s];
          orr = mfma64(A.x, B.x, orr);
          oii = mfma64(A.y, B.y, oii);
          ori = mfma64(A.x, B.y, ori);
          oir = mfma64(A.y, B.x, oir);
        }
        const mfma_d4 ur = orr - oii, ui = ori + oir;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int j = c0 + li;
          if (vD[v] && j >= fD[v]) S[rbD[v] + (j - fD[v])] = make_double2(ur[v], ui[v]);
        }
        // U[R,C] is read back by the later blocks of this row: the stores
        // complete and this CU's L1 drops lines loaded before them
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        __syncthreads();  // xs is reused
        continue;
      }
      // ---- the diagonal block: pivots row by row, the rest of the block
      // updated after each (lane l: row l & 15, columns (l >> 4) + 4q)
      for (int k = 0; k < nb; k++) {
        const double2 piv = crecip_smith(xs[k][k]);
        if (l == 0) pv[k] = piv;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int i = li, j = lk + 4 * q;
          if (j > k && j <= i && i < nb) xs[i][j] = csub(xs[i][j], cmul(xs[i][k], cmul(xs[j][k], piv)));
        }
        __syncthreads();
      }
      // l_ij = u_ij / d_j inside the block
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = li, j = lk + 4 * q;
        if (j < i && i < nb) xs[i][j] = cmul(xs[i][j], pv[j]);
      }
      __syncthreads();
    }

    // ---- block row R finished: scale its profile entries, store the
    // in-block part and the pivots
    for (int r = 0; r < nb; r++) {
      const int i = r0 + r, f = idx1[i], rb = rowbeg[i];
      for (int t = l; t <= i - f; t += 64) {
        const int j = f + t;
        double2 v;
        if (j < r0)
          v = cmul(S[rb + t], dinv[j]);
        else if (j < i)
          v = xs[r][j - r0];
        else
          v = pv[r];
        S[rb + t] = v;
      }
    }
    if (l < nb) dinv[r0 + l] = pv[l];
    // M_R = L_RR^-1 (unit lower): lane q < 16 forms column q,
    // (L^-1)_rq = -(L_rq + sum_{q<j<r} L_rj (L^-1)_jq)
    if (l < 16) {
      const int q = l;
      double2 x[16];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        double2 acc = cz();
        if (r > q && r < nb) {
          acc = xs[r][q];
#pragma unroll
          for (int j = 1; j < r; j++)
            if (j > q) acc = cfma(acc, xs[r][j], x[j]);
        }
        x[r] = r < q ? cz() : (r == q ? make_double2(1.0, 0.0) : csub(cz(), acc));
      }
      double2 *Mr = Mall + (long long)R * 256;
#pragma unroll
      for (int r = 0; r < 16; r++) Mr[r * 16 + q] = x[r];
    }
    // the finished rows, pivots and M_R are read by the later block rows
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
    __syncthreads();
  }
}

hipError_t launch_factor16(const FactorArgs &a, int nwork, hipStream_t st) {
  if (nwork <= 0) return hipSuccess;
  k_factor16<<<nwork, 64, 0, st>>>(a);
  return hipGetLastError();
}

}  // namespace fw
