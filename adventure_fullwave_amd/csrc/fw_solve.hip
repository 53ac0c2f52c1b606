// fw_solve.hip - batched subdomain Dirichlet solve (the Schur-step kernel).
//
// One 256-thread workgroup per subdomain (hddm_solver.h:461-693,
// advlas_skyline.c:134-167):
//   1. load: v_I = f_I (first/last step) - A_IG v_G, v_G = interface values
//      (SetBC, aij_matrix_func.h:60-136; renewal_bc_sp hddm_solver_sp.h:58-69)
//   2. forward sweep  v_i -= sum_j L_ij v_j      (row blocks of <= BMAX rows)
//   3. v_i *= 1/d_i
//   4. backward sweep v_j -= L_ij v_i, i descending
//   5. y_G = (A x)_G (- f_G)   (MxV_TMP_AIJ on the original matrix, rows of G)
//
// The skyline is streamed once per sweep in contiguous row blocks of at most
// CAP entries.  All index metadata of the subdomain (first columns, row
// starts, block starts) is staged in LDS up front, so the only global
// traffic of a sweep is the skyline itself, prefetched two blocks ahead in
// two register sets while the current block is processed from LDS.
#include "fw_kernels.h"
#include "fw_device.h"

namespace fw {

template <int PF>
__device__ __forceinline__ void issue_block(const double2 *__restrict__ fac, int e0, int e1, double2 (&p)[PF]) {
#pragma unroll
  for (int k = 0; k < PF; k++) {
    const int e = e0 + (int)threadIdx.x + k * SOLVE_THREADS;
    p[k] = e < e1 ? fac[e] : cz();
  }
}

template <int PF>
__device__ __forceinline__ void stage_block(double2 *stg, int cnt, const double2 (&p)[PF]) {
#pragma unroll
  for (int k = 0; k < PF; k++) {
    const int e = (int)threadIdx.x + k * SOLVE_THREADS;
    if (e < cnt) stg[e] = p[k];
  }
}

struct SolveLds {
  double2 *v, *vg, *stg;
  int *idx1, *rb, *blk;
  double2 *part;
};

// forward block b: rows [i0,i1) with skyline entries [e0,e1)
__device__ __forceinline__ void fwd_block(const SolveLds &L, const double2 *__restrict__ fac, int nblk, int b,
                                          double2 (&p)[PFK]) {
  const int tid = threadIdx.x;
  const int i0 = L.blk[b], i1 = L.blk[b + 1];
  const int e0 = L.rb[i0], e1 = L.rb[i1];
  const int nb = i1 - i0;
  __syncthreads();  // the previous block is done with stg, part and v
  stage_block(L.stg, e1 - e0, p);
  __syncthreads();
  if (b + 2 < nblk) issue_block(fac, L.rb[L.blk[b + 2]], L.rb[L.blk[b + 3]], p);
  // A: columns before the block, 16 lanes per row
  {
    const int r = tid >> 4, l = tid & 15;
    double2 acc = cz();
    if (r < nb) {
      const int i = i0 + r;
      const int fi = L.idx1[i], rb = L.rb[i] - e0;
      for (int j = fi + l; j < i0; j += 16) acc = cfma(acc, L.stg[rb + (j - fi)], L.v[j]);
    }
    acc = group_csum<16>(acc);
    if (r < nb && l == 0) L.part[r] = acc;
  }
  __syncthreads();
  // B: the block's own lower triangle, one lane per row (wave 0)
  if (tid < 64) {
    const int r = tid;
    double2 s = cz();
    int fi = 0, rb = 0;
    if (r < nb) {
      s = csub(L.v[i0 + r], L.part[r]);
      fi = L.idx1[i0 + r];
      rb = L.rb[i0 + r] - e0;
    }
    for (int q = 0; q < nb; q++) {
      const double2 xq = readlane_c(s, q);
      const int j = i0 + q;
      if (r > q && r < nb && j >= fi) s = cfms(s, L.stg[rb + (j - fi)], xq);
    }
    if (r < nb) L.v[i0 + r] = s;
  }
}

// backward block b (rows descending): triangle, then columns before the block
__device__ __forceinline__ void bwd_block(const SolveLds &L, const double2 *__restrict__ fac, int b,
                                          double2 (&p)[PFK]) {
  const int tid = threadIdx.x;
  const int i0 = L.blk[b], i1 = L.blk[b + 1];
  const int e0 = L.rb[i0], e1 = L.rb[i1];
  const int nb = i1 - i0;
  __syncthreads();
  stage_block(L.stg, e1 - e0, p);
  __syncthreads();
  if (b >= 2) issue_block(fac, L.rb[L.blk[b - 2]], L.rb[L.blk[b - 1]], p);
  // A: triangle, one lane per column q (wave 0)
  if (tid < 64) {
    const int q = tid;
    double2 x = q < nb ? L.v[i0 + q] : cz();
    for (int r = nb - 1; r >= 0; r--) {
      const double2 xr = readlane_c(x, r);
      const int fr = L.idx1[i0 + r];
      if (q < r && i0 + q >= fr) x = cfms(x, L.stg[L.rb[i0 + r] - e0 + (i0 + q - fr)], xr);
    }
    if (q < nb) L.v[i0 + q] = x;
  }
  __syncthreads();
  // B: columns before the block, all rows of the block
  int jlo = i0;
  for (int r = 0; r < nb; r++) jlo = min(jlo, L.idx1[i0 + r]);
  for (int j = jlo + tid; j < i0; j += SOLVE_THREADS) {
    double2 acc = cz();
    for (int r = 0; r < nb; r++) {
      const int f = L.idx1[i0 + r];
      if (j >= f) acc = cfma(acc, L.stg[L.rb[i0 + r] - e0 + (j - f)], L.v[i0 + r]);
    }
    L.v[j] = csub(L.v[j], acc);
  }
}

__global__ __launch_bounds__(SOLVE_THREADS) void k_solve(SolveArgs a) {
  if (a.mode == MODE_ITER && a.done && *a.done) return;
  const int w = a.order ? a.order[a.base + blockIdx.x] : a.base + blockIdx.x;
  const SubMeta m = a.meta[w];
  extern __shared__ double2 lds[];
  __shared__ double2 part[BMAX];
  SolveLds L;
  L.v = lds;                          // [nmax]
  L.vg = L.v + a.nmax;                // [ngmax]
  L.stg = L.vg + a.ngmax;             // [CAP]
  L.idx1 = (int *)(L.stg + CAP);      // [nmax]
  L.rb = L.idx1 + a.nmax;             // [nmax+1]
  L.blk = L.rb + a.nmax + 1;          // [nblkmax+1]
  L.part = part;
  const int tid = threadIdx.x;
  const int n = m.n;
  const double2 *fac = a.fac + m.fac0;

  // ---- 0. metadata and interface values --------------------------------------
  for (int i = tid; i < n; i += SOLVE_THREADS) {
    L.idx1[i] = a.idx1[m.dof0 + i];
    L.rb[i] = a.rowbeg[m.dof0 + i];
  }
  if (tid == 0) L.rb[n] = (int)m.nlen;
  for (int b = tid; b <= m.nblk; b += SOLVE_THREADS) L.blk[b] = a.blk_row[m.blk0 + b];
  for (int k = tid; k < m.ng; k += SOLVE_THREADS)
    L.vg[k] = (a.mode == MODE_FIRST || a.mode == MODE_RHS) ? cz() : a.P[a.g_gid[m.g0 + k]];
  __syncthreads();

  // first two skyline blocks in flight while the load vector is built
  double2 pa[PFK], pb[PFK];
  if (m.nblk > 0) issue_block(fac, L.rb[L.blk[0]], L.rb[L.blk[1]], pa);
  if (m.nblk > 1) issue_block(fac, L.rb[L.blk[1]], L.rb[L.blk[2]], pb);

  // ---- 1. load vector ------------------------------------------------------------
  for (int i = tid; i < n; i += SOLVE_THREADS) {
    const int d = m.dof0 + i;
    double2 b;
    if (a.mode == MODE_RHS) {
      b = a.X[d];
    } else {
      const int8_t fl = a.bcflag[d];
      if (fl == 1) {
        b = L.vg[a.kidx[d]];
      } else if (fl == 2) {
        b = cz();
      } else {
        b = (a.mode == MODE_ITER) ? cz() : a.f[d];
        for (int c = a.cp_ptr[d]; c < a.cp_ptr[d + 1]; c++) b = cfms(b, a.cp_val[c], L.vg[a.cp_k[c]]);
      }
    }
    L.v[i] = b;
  }

  // ---- 2. forward sweep ---------------------------------------------------------------
  for (int b = 0; b < m.nblk; b += 2) {
    fwd_block(L, fac, m.nblk, b, pa);
    if (b + 1 < m.nblk) fwd_block(L, fac, m.nblk, b + 1, pb);
  }
  // first two backward blocks in flight during the diagonal pass
  const int nb0 = m.nblk - 1;
  if (nb0 >= 0) {
    // block nb0 goes to the set it would use in the backward loop below
    if ((nb0 & 1) == 0) issue_block(fac, L.rb[L.blk[nb0]], L.rb[L.blk[nb0 + 1]], pa);
    else issue_block(fac, L.rb[L.blk[nb0]], L.rb[L.blk[nb0 + 1]], pb);
  }
  if (nb0 >= 1) {
    if (((nb0 - 1) & 1) == 0) issue_block(fac, L.rb[L.blk[nb0 - 1]], L.rb[L.blk[nb0]], pa);
    else issue_block(fac, L.rb[L.blk[nb0 - 1]], L.rb[L.blk[nb0]], pb);
  }
  __syncthreads();

  // ---- 3. diagonal ------------------------------------------------------------------------
  for (int i = tid; i < n; i += SOLVE_THREADS) L.v[i] = cmul(L.v[i], a.dinv[m.dof0 + i]);

  // ---- 4. backward sweep: even blocks use pa, odd blocks pb -------------------------------
  for (int b = nb0; b >= 0; b--) {
    if ((b & 1) == 0) bwd_block(L, fac, b, pa);
    else bwd_block(L, fac, b, pb);
  }
  __syncthreads();

  // ---- 5. outputs --------------------------------------------------------------------------
  if (a.mode == MODE_LAST || a.mode == MODE_RHS)
    for (int i = tid; i < n; i += SOLVE_THREADS) a.X[m.dof0 + i] = L.v[i];
  if (a.mode != MODE_RHS) {
    for (int k = tid; k < m.ng; k += SOLVE_THREADS) {
      const int sl = m.g0 + k;
      double2 y = cz();
      for (int c = a.gr_ptr[sl]; c < a.gr_ptr[sl + 1]; c++) y = cfma(y, a.gr_val[c], L.v[a.gr_col[c]]);
      if (a.mode != MODE_ITER) y = csub(y, a.f[m.dof0 + a.g_nfid[sl]]);
      a.yslot[sl] = y;
    }
  }
}

size_t solve_lds_bytes(int nmax, int ngmax, int nblkmax) {
  return sizeof(double2) * ((size_t)nmax + ngmax + CAP) + sizeof(int) * ((size_t)2 * nmax + 1 + nblkmax + 1);
}

hipError_t launch_solve(const SolveArgs &a, int nwork, hipStream_t st) {
  if (nwork <= 0) return hipSuccess;
  k_solve<<<nwork, SOLVE_THREADS, solve_lds_bytes(a.nmax, a.ngmax, a.nblkmax), st>>>(a);
  return hipGetLastError();
}

}  // namespace fw
