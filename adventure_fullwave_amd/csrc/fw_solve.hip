// fw_solve.hip - batched subdomain Dirichlet solve (the Schur-step kernel).
//
// One wavefront (64 lanes) per subdomain (hddm_solver.h:461-693,
// advlas_skyline.c:134-167):
//   1. load: v_I = f_I (first/last step) - A_IG v_G, v_G = interface values
//      (SetBC, aij_matrix_func.h:60-136; renewal_bc_sp hddm_solver_sp.h:58-69)
//   2. forward sweep  v_i -= sum_j L_ij v_j      (row blocks of <= BMAX rows)
//   3. v_i *= 1/d_i
//   4. backward sweep v_j -= L_ij v_i, i descending
//   5. y_G = (A x)_G (- f_G)   (MxV_TMP_AIJ on the original matrix, rows of G)
//
// The skyline is streamed once per sweep in contiguous row blocks of at most
// 64*K entries (K = 8 or 16 loads per lane; the host keeps a row-block
// partition for each K and picks K per launch kind), two blocks in flight per
// wavefront (register sets pa/pb), each staged in LDS when its turn comes.  A sweep is a chain of
// dependent rows, so the kernel is organised around that chain: no
// workgroup barriers, block-row metadata in registers (broadcast with
// v_readlane), the triangle of each block resolved by one lane per row with
// its operands preloaded into registers.
#include <algorithm>

#include "fw_kernels.h"
#include "fw_device.h"

namespace fw {

constexpr int WL = 64;  // lanes

__device__ __forceinline__ int readlane_i(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }

// one prefetched block per lane: plain doubles (a POD the compiler keeps in
// registers; the HIP vector type's unions make it fall back to scratch)
struct Pf2 {
  double x, y;
};
// a row block: rows [i0, i1), skyline entries [e0, e1)
struct BlkDesc {
  int i0, i1, e0, e1;
};
template <int K>
struct PfSet {
  Pf2 e[K];
  Pf2 m[MINV / 64];  // the block's inverted diagonal triangle
  BlkDesc q;         // the block itself (read with the addresses of its loads)
};

// Branch-free block loads: addresses past the block are clamped to its last
// entry (same cache line, no extra traffic), so the loads carry no control
// flow and the compiler can keep the other register set's loads in flight
// (a counted vmcnt wait instead of a full drain).
template <int K>
__device__ __forceinline__ void issue_block(const double2 *__restrict__ fac, const BlkDesc q,
                                            const double2 *__restrict__ minv, PfSet<K> &p) {
  const int lane = threadIdx.x & (WL - 1);
  const int e0 = q.e0, last = q.e1 - 1;
  p.q = q;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double2 t = fac[min(e0 + lane + k * WL, last)];
    p.e[k].x = t.x;
    p.e[k].y = t.y;
  }
#pragma unroll
  for (int k = 0; k < MINV / WL; k++) {
    const double2 t = minv[lane + k * WL];
    p.m[k].x = t.x;
    p.m[k].y = t.y;
  }
}

// Unconditional staging: entries past the block land in unused LDS slots.
template <int K>
__device__ __forceinline__ void stage_block(double2 *stg, double2 *mst, const PfSet<K> &p) {
  const int lane = threadIdx.x & (WL - 1);
#pragma unroll
  for (int k = 0; k < K; k++) stg[lane + k * WL] = make_double2(p.e[k].x, p.e[k].y);
#pragma unroll
  for (int k = 0; k < MINV / WL; k++) mst[lane + k * WL] = make_double2(p.m[k].x, p.m[k].y);
}

// The workgroup is a single wavefront: LDS traffic of one wave is processed
// in order, so a compiler ordering point is all a phase change needs (no
// s_barrier and, above all, no memory fence that would drain the loads in
// flight).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// A team is the group of threads that solves one subdomain: a workgroup of
// NT threads, or (NT = 64) one wave, alone in its workgroup or one of the
// waves of a persistent workgroup that solve subdomains each on their own.
template <int NT>
__device__ __forceinline__ int team_tid() {
  return NT == WL ? (int)(threadIdx.x & (WL - 1)) : (int)threadIdx.x;
}
// ordering point of a single-wave team: a barrier when the wave is the
// workgroup, else a wave barrier (LDS traffic of one wave is processed in
// order, so the other waves need not wait)
template <bool TEAM>
__device__ __forceinline__ void team_sync() {
  if (TEAM)
    wave_sync();
  else
    __syncthreads();
}

struct SolveLds {
  double2 *v, *stg, *mst;       // LDS: vector, staged block, staged inverse triangle
  int *idx1, *rb, *blk;         // LDS: row metadata, block rows
  BlkDesc *bt;                  // LDS: every row block's rows and entries
  const double2 *minv;          // global: inverse triangles of this subdomain's blocks
};
// inverse triangle of block b (packed rows, MINV entries per block)
__device__ __forceinline__ const double2 *minv_of(const SolveLds &L, int b) { return L.minv + (long long)b * MINV; }

// Backward column updates of a block: v_j -= sum_r L_rj x_r for the columns
// j in [jlo, i0) (thread t takes j = jlo + t, jlo + t + NT, ...), rows in
// ascending order.  Row r's first column, base and solution value are on
// lane xl * r of the wave (fR, rbR, xv).  R (4, 8 or 16, at least the
// block's rows) rows are summed without a branch, so their R LDS reads are
// issued together: a row past the block has fR = i0 > j and adds a valid
// entry times zero.
template <int R, int NT>
__device__ __forceinline__ void col_updates(const SolveLds &L, int jlo, int i0, int fR, int rbR, double2 xv,
                                            int xl) {
  for (int j = jlo + team_tid<NT>(); j < i0; j += NT) {
    double2 acc = cz();
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int f_u = readlane_i(fR, r), rb_u = readlane_i(rbR, r);
      const double2 xr = readlane_c(xv, xl * r);
      const bool in = j >= f_u;
      acc = cfma(acc, L.stg[in ? rb_u + j : 0], in ? xr : cz());
    }
    L.v[j] = csub(L.v[j], acc);
  }
}
template <int NT>
__device__ __forceinline__ void col_updates_nb(const SolveLds &L, int nb, int jlo, int i0, int fR, int rbR,
                                               double2 xv, int xl) {
  if (nb <= 4)
    col_updates<4, NT>(L, jlo, i0, fR, rbR, xv, xl);
  else if (nb <= 8)
    col_updates<8, NT>(L, jlo, i0, fR, rbR, xv, xl);
  else
    col_updates<BMAX, NT>(L, jlo, i0, fR, rbR, xv, xl);
}

// forward block b: rows [i0,i1) with skyline entries [e0,e1)
template <int K>
__device__ __forceinline__ void fwd_block(const SolveLds &L, const double2 *__restrict__ fac, int nblk, int b,
                                          PfSet<K> &p) {
  const int lane = threadIdx.x & (WL - 1);
  const int i0 = p.q.i0, i1 = p.q.i1;
  const int e0 = p.q.e0;
  const int nb = i1 - i0;
  wave_sync();
  stage_block(L.stg, L.mst, p);
  wave_sync();
  {
    const int bn = min(b + 2, nblk - 1);  // unconditional: see issue_block
    issue_block(fac, L.bt[bn], minv_of(L, bn), p);
  }

  // A: columns before the block.  Lr lanes per row, Lr = 64 / nb rounded
  // down to a power of two (at least 4), so a block of a few wide rows keeps
  // every lane busy.  A lane takes every Lr-th column of its row, four at a
  // time: the batch's eight LDS reads are issued together (addresses
  // clamped, terms past the row multiplied by zero), then added in column
  // order into two partial sums; a sum over the row's lanes adds them.
  const int lsh = nb <= 1 ? 6 : nb <= 2 ? 5 : nb <= 4 ? 4 : nb <= 8 ? 3 : 2;
  const int Lr = 1 << lsh, ra = lane >> lsh, la = lane & (Lr - 1);
  double2 acc0 = cz(), acc1 = cz();
  if (ra < nb) {
    const int fi = L.idx1[i0 + ra];
    const int rs = L.rb[i0 + ra] - e0;  // the row's first staged entry
    const int rbase = rs - fi;
    for (int j0 = fi + la; j0 < i0; j0 += 4 * Lr) {
      double2 sv[4], vv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = j0 + u * Lr;
        const bool in = j < i0;
        sv[u] = L.stg[in ? rbase + j : rs];
        vv[u] = L.v[in ? j : 0];
      }
#pragma unroll
      for (int u = 0; u < 4; u += 2) {
        acc0 = cfma(acc0, sv[u], j0 + u * Lr < i0 ? vv[u] : cz());
        acc1 = cfma(acc1, sv[u + 1], j0 + (u + 1) * Lr < i0 ? vv[u + 1] : cz());
      }
    }
  }
  const double2 acc = group_csum(cadd(acc0, acc1), lsh);
  if (la == 0 && ra < nb) L.v[i0 + ra] = csub(L.v[i0 + ra], acc);
  wave_sync();

  // B: the block's own triangle.  With s = v_B - (columns before the block)
  // and M = L_BB^-1 (unit lower, precomputed at load), x_r = s_r +
  // sum_{q<r} M_rq s_q: four lanes per row, one term of every fourth column
  // each, then a quad sum (no chain of dependent steps)
  const int r = lane >> 2, l4 = lane & 3;
  const bool own = l4 == 0 && r < nb;
  const double2 s = own ? L.v[i0 + r] : cz();
  double2 t = cz();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = l4 + 4 * k;
    const bool use = q < r && r < nb;
    const double2 mq = L.mst[use ? (r * (r - 1)) / 2 + q : 0];
    const double2 sq = L.v[use ? i0 + q : 0];
    t = cfma(t, use ? mq : cz(), sq);
  }
  t = group_csum(t, 2);
  if (own) L.v[i0 + r] = cadd(s, t);
}

// backward block b (rows descending): triangle, then columns before the block
template <int K>
__device__ __forceinline__ void bwd_block(const SolveLds &L, const double2 *__restrict__ fac, int b,
                                          PfSet<K> &p) {
  const int lane = threadIdx.x & (WL - 1);
  const int i0 = p.q.i0, i1 = p.q.i1;
  const int e0 = p.q.e0;
  const int nb = i1 - i0;
  wave_sync();
  stage_block(L.stg, L.mst, p);
  wave_sync();
  {
    const int bn = max(b - 2, 0);
    issue_block(fac, L.bt[bn], minv_of(L, bn), p);
  }

  // row metadata of the block: lane r holds row r's first column and base
  int fR = i0, rbR = 0;
  if (lane < nb) {
    fR = L.idx1[i0 + lane];
    rbR = L.rb[i0 + lane] - e0 - fR;
  }
  // A: triangle through M^T (L_BB^T x = v_B): x_q = v_q + sum_{r>q} M_rq v_r,
  // column q on lanes 4q..4q+3, one term of every fourth row each
  const int qc = lane >> 2, la = lane & 3;
  double2 t = cz();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int r = qc + 1 + la + 4 * k;
    const bool use = r < nb;
    const double2 mr = L.mst[use ? (r * (r - 1)) / 2 + qc : 0];
    const double2 vr = L.v[use ? i0 + r : 0];
    t = cfma(t, use ? mr : cz(), vr);
  }
  t = group_csum(t, 2);
  const bool own = la == 0 && qc < nb;
  double2 x = cz();
  if (own) x = cadd(L.v[i0 + qc], t);
  wave_sync();
  if (own) L.v[i0 + qc] = x;

  // B: every column before the block receives all rows of the block (row r's
  // value on lane 4r)
  int jlo = fR;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) jlo = min(jlo, __shfl_xor(jlo, o, 64));
  // (branch-free: a column outside row r's profile reads a valid entry and
  // multiplies it by zero, so the 16 LDS reads pipeline)
  col_updates_nb<WL>(L, nb, jlo, i0, fR, rbR, x, 4);
}

// the value arrays of system sys0 + blockIdx.y (index arrays are shared)
__device__ __forceinline__ void to_system(SolveArgs &a) {
  const long long s = a.sys0 + (long long)blockIdx.y;
  if (s == 0) return;
  a.fac += s * a.ss.nlen;
  a.dinv += s * a.ss.dof;
  a.f += s * a.ss.dof;
  a.cp_val += s * a.ss.cpl;
  a.gr_val += s * a.ss.grow;
  a.P += s * a.ss.g;
  if (a.Z) a.Z += s * a.ss.g;
  a.sc += s;
  if (a.done) a.done = &a.sc->done;
  a.yslot += s * a.ss.slot;
  a.X += s * a.ss.dof;
  if (a.smat) a.smat += s * a.ss.smat;
  if (a.pqpart) a.pqpart += s * (long long)a.npq;
  a.minv += s * a.ss.minv;
}

// ---------------------------------------------------------------------------
// Parts of the Schur step over a subdomain's DOFs, shared by both forms of
// the kernel (NT threads, thread t takes DOFs t, t + NT, ...)

// the subdomain's work item: (subdomain, unit vector) of MODE_UNIT, else
// the subdomain; false: nothing to do
__device__ __forceinline__ bool solve_item(const SolveArgs &a, int &w, int &uk) {
  uk = a.unit_k;
  if (a.mode == MODE_UNIT && a.unit_list) {
    const int2 pr = a.unit_list[a.base + blockIdx.x];  // (subdomain, unit vector); -1: padding
    if (pr.x < 0) return false;
    w = pr.x;
    uk = pr.y;
  } else {
    w = a.order ? a.order[a.base + blockIdx.x] : a.base + blockIdx.x;
  }
  return true;
}

// LDS carve-up: v[nmax], stg[cap], mst[nmst], bt[nblkmax+1] (16-byte
// entries), idx1[nmax], rb[nmax+1], blk[nblkmax+1]
__device__ __forceinline__ SolveLds solve_lds(double2 *lds, int nmax, int nblkmax, int cap, int nmst) {
  SolveLds L;
  L.v = lds;
  L.stg = L.v + nmax;
  L.mst = L.stg + cap;
  L.bt = (BlkDesc *)(L.mst + nmst);
  L.idx1 = (int *)(L.bt + nblkmax + 1);
  L.rb = L.idx1 + nmax;
  L.blk = L.rb + nmax + 1;
  return L;
}
__device__ __forceinline__ SolveLds solve_lds(const SolveArgs &a, double2 *lds, int cap, int nmst) {
  return solve_lds(lds, a.nmax, a.nblkmax, cap, nmst);
}
// the row-block table from the block rows and row starts in LDS (call
// after a barrier that follows load_meta)
template <int NT>
__device__ __forceinline__ void load_blocks(const SolveLds &L, int nblk) {
  for (int b = team_tid<NT>(); b < nblk; b += NT) {
    const int i0 = L.blk[b], i1 = L.blk[b + 1];
    L.bt[b] = BlkDesc{i0, i1, L.rb[i0], L.rb[i1]};
  }
}

template <int NT>
__device__ __forceinline__ void load_meta(const SolveArgs &a, const SubMeta &m, const SolveLds &L, int nblk,
                                          int blk0) {
  const int t = team_tid<NT>();
  for (int i = t; i < m.n; i += NT) {
    L.idx1[i] = a.idx1[m.dof0 + i];
    L.rb[i] = a.rowbeg[m.dof0 + i];
  }
  if (t == 0) L.rb[m.n] = (int)m.nlen;
  for (int b = t; b <= nblk; b += NT) L.blk[b] = a.blk_row[blk0 + b];
}

// the load vector (first/last step: f_I - A_IG v_G; iteration: -A_IG p_G;
// unit vector; right-hand side) into v.  pvs: LDS for cap interface values
// that is free until the sweeps start (the staging buffer): when the
// subdomain's interface values fit, they are gathered there once, so the
// coupling loop reads them from LDS instead of gathering per coupling.
template <int NT>
__device__ __forceinline__ void load_rhs(const SolveArgs &a, const SubMeta &m, double2 *v, int uk, double2 *pvs,
                                         int cap) {
  const bool zero_g = (a.mode == MODE_FIRST || a.mode == MODE_RHS);
  // interface values: P, or the next search direction z + beta p while the
  // interface iteration runs (formed here, not stored, see k_iter_ab)
  const double2 beta = (a.mode == MODE_ITER && a.Z) ? a.sc->beta : cz();
  auto pval = [&](int gid) -> double2 {
    const double2 p = a.P[gid];
    return a.Z ? zbp(a.Z[gid], beta, p) : p;
  };
  // (uniform over the team: the same subdomain and mode)
  const bool staged = (a.mode == MODE_ITER || a.mode == MODE_LAST) && m.ng <= cap;
  if (staged) {
    for (int k = team_tid<NT>(); k < m.ng; k += NT) pvs[k] = pval(a.g_gid[m.g0 + k]);
    if (NT == WL)
      wave_sync();  // one wave: its LDS operations complete in order
    else
      __syncthreads();
  }
  auto gval = [&](int k) -> double2 { return staged ? pvs[k] : pval(a.g_gid[m.g0 + k]); };
  for (int i = team_tid<NT>(); i < m.n; i += NT) {
    const int d = m.dof0 + i;
    double2 b;
    if (a.mode == MODE_RHS) {
      b = a.X[d];
    } else if (a.mode == MODE_UNIT) {
      const int8_t fl = a.bcflag[d];
      b = cz();
      if (fl == 1) {
        if (a.kidx[d] == uk) b.x = 1.0;
      } else if (fl == 0) {
        for (int c = a.cp_ptr[d]; c < a.cp_ptr[d + 1]; c++)
          if (a.cp_k[c] == uk) b = csub(b, a.cp_val[c]);
      }
    } else {
      const int8_t fl = a.bcflag[d];
      if (fl == 1) {
        b = zero_g ? cz() : gval(a.kidx[d]);
      } else if (fl == 2) {
        b = cz();
      } else {
        b = (a.mode == MODE_ITER) ? cz() : a.f[d];
        if (!zero_g)
          for (int c = a.cp_ptr[d]; c < a.cp_ptr[d + 1]; c++) b = cfms(b, a.cp_val[c], gval(a.cp_k[c]));
      }
    }
    v[i] = b;
  }
}

// outputs from the solution v: X (last step, right-hand side), a column of
// S_s (unit vector), else y_G = (A x)_G (- f_G) per slot; returns this
// thread's part of (p,q) (the interface iteration's sum over the
// subdomain's slots of p y: an interface row's solution value is p itself,
// its factor row being the identity and its column in L zero)
template <int NT>
__device__ __forceinline__ double2 write_outputs(const SolveArgs &a, const SubMeta &m, const double2 *v, int w,
                                                 int uk) {
  const int t = team_tid<NT>();
  double2 pq = cz();
  if (a.mode == MODE_LAST || a.mode == MODE_RHS)
    for (int i = t; i < m.n; i += NT) a.X[m.dof0 + i] = v[i];
  if (a.mode == MODE_UNIT) {
    // column unit_k of S_s: rows k >= unit_k into the lower-triangle tiles
    // (layout in fw_schur.hip), the diagonal halved
    const int j = uk, J = j >> 4, cj = j & 15;
    const int nbk = (m.ng + 15) >> 4;
    const bool hlf = smv_half(m.ng);
    double2 *dst = a.smat + a.sm_off[w];
    for (int k = j + t; k < m.ng; k += NT) {
      const int sl = m.g0 + k;
      double2 y = cz();
      for (int c = a.gr_ptr[sl]; c < a.gr_ptr[sl + 1]; c++) y = cfma(y, a.gr_val[c], v[a.gr_col[c]]);
      if (k == j) y = cscale(y, 0.5);
      dst[tile_pos(k >> 4, J, k & 15, cj, nbk, hlf)] = y;
    }
  } else if (a.mode != MODE_RHS) {
    for (int k = t; k < m.ng; k += NT) {
      const int sl = m.g0 + k;
      double2 y = cz();
      for (int c = a.gr_ptr[sl]; c < a.gr_ptr[sl + 1]; c++) y = cfma(y, a.gr_val[c], v[a.gr_col[c]]);
      if (a.mode != MODE_ITER) y = csub(y, a.f[m.dof0 + a.g_nfid[sl]]);
      a.yslot[sl] = y;
      pq = cadd(pq, cmul(v[a.g_nfid[sl]], y));
    }
  }
  return pq;
}

// The Schur step of subdomain w on one wave (K prefetch loads per lane: row
// blocks of at most 64 K entries), LDS carved in L.  TEAM: the wave is one
// of several in its workgroup, each on its own subdomain.  done0: the
// system is finished (the launch stops once the metadata loads are issued).
template <int K, bool TEAM>
__device__ __forceinline__ void solve1_body(const SolveArgs &a, const SubMeta &m, int w, int uk, SolveLds L,
                                            int done0) {
  const int lane = threadIdx.x & (WL - 1);
  const int n = m.n;
  const double2 *fac = a.fac + m.fac0;
  // row blocks of this prefetch depth (the host keeps both partitions)
  const int nblk = K == 8 ? m.nblk_s : m.nblk;
  const int blk0 = K == 8 ? m.blk0_s : m.blk0;
  L.minv = a.minv + (long long)blk0 * MINV;

  // ---- 0. metadata --------------------------------------------------------------
  load_meta<WL>(a, m, L, nblk, blk0);
  team_sync<TEAM>();
  if (done0) return;
  load_blocks<WL>(L, nblk);
  team_sync<TEAM>();

  // ---- 1. load vector --------------------------------------------------------------
  load_rhs<WL>(a, m, L.v, uk, L.stg, WL * K);  // the staging buffer is free until the forward sweep

  // ---- 2. forward sweep: the first two blocks go in flight once the load
  // vector's own global reads are complete ----------------------------------
  PfSet<K> pa, pb;
  // (the host pads every subdomain to an even number >= 2 of row blocks, so
  // the sweeps run over block pairs without conditional loads: a branch
  // around a load would make the compiler drain every load in flight)
  issue_block(fac, L.bt[0], minv_of(L, 0), pa);
  issue_block(fac, L.bt[1], minv_of(L, 1), pb);
  for (int b = 0; b < nblk; b += 2) {
    fwd_block(L, fac, nblk, b, pa);
    fwd_block(L, fac, nblk, b + 1, pb);
  }
  wave_sync();

  // ---- 3. diagonal (before the backward loads are issued: global loads
  // complete in issue order, so a later dinv load would drain them) ---------
  for (int i = lane; i < n; i += WL) L.v[i] = cmul(L.v[i], a.dinv[m.dof0 + i]);
  wave_sync();

  // the last two blocks go in flight for the backward sweep: the last block
  // in pa, the one before in pb, alternating from there (the register sets
  // are always named statically, so they stay in registers)
  const int nbl = nblk - 1;
  issue_block(fac, L.bt[nbl], minv_of(L, nbl), pa);
  issue_block(fac, L.bt[nbl - 1], minv_of(L, nbl - 1), pb);
  wave_sync();

  // ---- 4. backward sweep ---------------------------------------------------------------------
  for (int b = nbl; b >= 1; b -= 2) {
    bwd_block(L, fac, b, pa);
    bwd_block(L, fac, b - 1, pb);
  }
  wave_sync();

  // ---- 5. outputs ------------------------------------------------------------------------------
  double2 pq = write_outputs<WL>(a, m, L.v, w, uk);
  if (a.mode == MODE_ITER && a.pqpart) {
    pq = wave_csum(pq);
    if (lane == 0) a.pqpart[w] = pq;
  }
}

template <int K>  // prefetch loads per lane: row blocks of at most 64*K entries
__global__ __launch_bounds__(WL) __attribute__((amdgpu_waves_per_eu(2, K == 8 ? 3 : 2))) void k_solve(SolveArgs a) {
  constexpr int CAPK = WL * K;
  to_system(a);
  // (the status of an iteration launch is read below, once the metadata
  // loads are in flight)
  const bool check_done = a.mode == MODE_ITER && a.done;
  const int done0 = check_done ? *a.done : 0;
  int w = 0, uk = 0;
  if (!solve_item(a, w, uk)) return;
  const SubMeta m = a.meta[w];
  if (a.mode == MODE_UNIT && uk >= m.ng) return;
  extern __shared__ double2 lds[];
  solve1_body<K, false>(a, m, w, uk, solve_lds(a, lds, CAPK, MINV), done0);
}

// ---------------------------------------------------------------------------
// The same Schur step with two or four waves per subdomain (NT = 128 or 256
// threads).  A sweep stays a chain of row blocks, but the work of each block
// is spread over the workgroup: every thread stages and prefetches its share
// of the block (KQ loads each, blocks of at most NT*KQ entries: the
// partitions of the single-wave kernel's depths 16 and 8), the dot products
// of the forward block run with up to 64 lanes per row, the backward
// block's column updates over NT columns at a time; wave 0 resolves the
// 16x16 triangles through L_BB^-1 as the single-wave kernel does.  Workgroup barriers order the phases: forward,
// staged -> dot products -> triangle (the next block's barrier orders the
// triangle before its dot products; the inverse triangles are staged
// alternately in two slots, so the next block's staging does not touch the
// one wave 0 may still read); backward, staged -> triangle -> column updates
// -> (all updates done before the next block is staged).  Sums run in a
// fixed order, so the results do not depend on scheduling; the order differs
// from the single-wave kernel's (more lanes per row), so the two forms agree
// to rounding.

template <int KQ>
struct PfSet4 {
  Pf2 e[KQ];
  Pf2 m;      // one entry of the block's inverted diagonal triangle (threads < MINV)
  BlkDesc q;  // the block itself
};

template <int KQ, int NT>
__device__ __forceinline__ void issue_block4(const double2 *__restrict__ fac, const BlkDesc q,
                                             const double2 *__restrict__ minv, PfSet4<KQ> &p) {
  const int t = threadIdx.x;
  const int e0 = q.e0, last = q.e1 - 1;
  p.q = q;
#pragma unroll
  for (int k = 0; k < KQ; k++) {
    const double2 v = fac[min(e0 + t + k * NT, last)];
    p.e[k].x = v.x;
    p.e[k].y = v.y;
  }
  const double2 v = minv[t & (MINV - 1)];  // threads >= MINV load a duplicate (same lines)
  p.m.x = v.x;
  p.m.y = v.y;
}

template <int KQ, int NT>
__device__ __forceinline__ void stage_block4(double2 *stg, double2 *mst, const PfSet4<KQ> &p) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < KQ; k++) stg[t + k * NT] = make_double2(p.e[k].x, p.e[k].y);
  if (t < MINV) mst[t] = make_double2(p.m.x, p.m.y);
}

template <int KQ, int NT>
__device__ __forceinline__ void fwd_block4(const SolveLds &L, const double2 *__restrict__ fac, int nblk, int b,
                                           PfSet4<KQ> &p) {
  const int t = threadIdx.x, lane = t & (WL - 1);
  const int i0 = p.q.i0, i1 = p.q.i1;
  const int e0 = p.q.e0;
  const int nb = i1 - i0;
  double2 *mst = L.mst + (b & 1) * MINV;
  stage_block4<KQ, NT>(L.stg, mst, p);
  __syncthreads();
  {
    const int bn = min(b + 2, nblk - 1);  // unconditional: see issue_block
    issue_block4<KQ, NT>(fac, L.bt[bn], minv_of(L, bn), p);
  }

  // A: columns before the block, Lr lanes per row: NT / nb rounded down to
  // a power of two, at most a wave (256 threads: 64 for blocks of up to four
  // rows, 32 up to eight, else 16), in batches of four columns per lane
  constexpr int LOG2NT = NT == 128 ? 7 : 8;
  const int lgnb = nb <= 1 ? 0 : nb <= 2 ? 1 : nb <= 4 ? 2 : nb <= 8 ? 3 : 4;
  const int lsh = min(6, LOG2NT - lgnb);
  const int Lr = 1 << lsh, ra = t >> lsh, la = t & (Lr - 1);
  double2 acc0 = cz(), acc1 = cz();
  if (ra < nb) {
    const int fi = L.idx1[i0 + ra];
    const int rs = L.rb[i0 + ra] - e0;
    const int rbase = rs - fi;
    for (int j0 = fi + la; j0 < i0; j0 += 4 * Lr) {
      double2 sv[4], vv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = j0 + u * Lr;
        const bool in = j < i0;
        sv[u] = L.stg[in ? rbase + j : rs];
        vv[u] = L.v[in ? j : 0];
      }
#pragma unroll
      for (int u = 0; u < 4; u += 2) {
        acc0 = cfma(acc0, sv[u], j0 + u * Lr < i0 ? vv[u] : cz());
        acc1 = cfma(acc1, sv[u + 1], j0 + (u + 1) * Lr < i0 ? vv[u + 1] : cz());
      }
    }
  }
  const double2 acc = group_csum(cadd(acc0, acc1), lsh);
  if (la == 0 && ra < nb) L.v[i0 + ra] = csub(L.v[i0 + ra], acc);
  __syncthreads();

  // B: the block's triangle on wave 0 (as in fwd_block)
  if (t < WL) {
    const int r = lane >> 2, l4 = lane & 3;
    const bool own = l4 == 0 && r < nb;
    const double2 s = own ? L.v[i0 + r] : cz();
    double2 tt = cz();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int q = l4 + 4 * k;
      const bool use = q < r && r < nb;
      const double2 mq = mst[use ? (r * (r - 1)) / 2 + q : 0];
      const double2 sq = L.v[use ? i0 + q : 0];
      tt = cfma(tt, use ? mq : cz(), sq);
    }
    tt = group_csum(tt, 2);
    if (own) L.v[i0 + r] = cadd(s, tt);
  }
}

template <int KQ, int NT>
__device__ __forceinline__ void bwd_block4(const SolveLds &L, const double2 *__restrict__ fac, int b,
                                           PfSet4<KQ> &p) {
  const int t = threadIdx.x, lane = t & (WL - 1);
  const int i0 = p.q.i0, i1 = p.q.i1;
  const int e0 = p.q.e0;
  const int nb = i1 - i0;
  double2 *mst = L.mst + (b & 1) * MINV;
  stage_block4<KQ, NT>(L.stg, mst, p);
  __syncthreads();
  {
    const int bn = max(b - 2, 0);
    issue_block4<KQ, NT>(fac, L.bt[bn], minv_of(L, bn), p);
  }
  // row metadata of the block, in every wave: lane r holds row r's first
  // column and base
  int fR = i0, rbR = 0;
  if (lane < nb) {
    fR = L.idx1[i0 + lane];
    rbR = L.rb[i0 + lane] - e0 - fR;
  }
  // A: triangle through M^T on wave 0 (as in bwd_block)
  if (t < WL) {
    const int qc = lane >> 2, l4 = lane & 3;
    double2 tt = cz();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int r = qc + 1 + l4 + 4 * k;
      const bool use = r < nb;
      const double2 mr = mst[use ? (r * (r - 1)) / 2 + qc : 0];
      const double2 vr = L.v[use ? i0 + r : 0];
      tt = cfma(tt, use ? mr : cz(), vr);
    }
    tt = group_csum(tt, 2);
    const bool own = l4 == 0 && qc < nb;
    double2 x = cz();
    if (own) x = cadd(L.v[i0 + qc], tt);
    wave_sync();
    if (own) L.v[i0 + qc] = x;
  }
  __syncthreads();

  // B: every column before the block receives all rows of the block (row
  // r's solution value on lane r of every wave)
  const double2 xR = lane < nb ? L.v[i0 + lane] : cz();
  int jlo = fR;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) jlo = min(jlo, __shfl_xor(jlo, o, 64));
  col_updates_nb<NT>(L, nb, jlo, i0, fR, rbR, xR, 1);
  __syncthreads();
}

// The Schur step of subdomain w on a workgroup of NT threads (KQ prefetch
// loads per thread: row blocks of at most NT KQ entries, the partition of
// the single-wave kernel with blocks of that size), LDS carved in L.
template <int KQ, int NT>
__device__ __forceinline__ void solve4_body(const SolveArgs &a, const SubMeta &m, int w, int uk, SolveLds L,
                                            int done0) {
  constexpr int CAPQ = NT * KQ;
  const int t = threadIdx.x;
  const int n = m.n;
  const double2 *fac = a.fac + m.fac0;
  const int nblk = CAPQ == 512 ? m.nblk_s : CAPQ == 2048 ? m.nblk_w : m.nblk;
  const int blk0 = CAPQ == 512 ? m.blk0_s : CAPQ == 2048 ? m.blk0_w : m.blk0;
  L.minv = a.minv + (long long)blk0 * MINV;

  load_meta<NT>(a, m, L, nblk, blk0);
  __syncthreads();
  if (done0) return;
  load_blocks<NT>(L, nblk);
  __syncthreads();
  load_rhs<NT>(a, m, L.v, uk, L.stg, CAPQ);  // the staging buffer is free until the forward sweep
  PfSet4<KQ> pa, pb;
  issue_block4<KQ, NT>(fac, L.bt[0], minv_of(L, 0), pa);
  issue_block4<KQ, NT>(fac, L.bt[1], minv_of(L, 1), pb);
  __syncthreads();  // the load vector is complete
  for (int b = 0; b < nblk; b += 2) {
    fwd_block4<KQ, NT>(L, fac, nblk, b, pa);
    fwd_block4<KQ, NT>(L, fac, nblk, b + 1, pb);
  }
  __syncthreads();  // the last triangle is complete
  for (int i = t; i < n; i += NT) L.v[i] = cmul(L.v[i], a.dinv[m.dof0 + i]);
  const int nbl = nblk - 1;
  issue_block4<KQ, NT>(fac, L.bt[nbl], minv_of(L, nbl), pa);
  issue_block4<KQ, NT>(fac, L.bt[nbl - 1], minv_of(L, nbl - 1), pb);
  __syncthreads();
  for (int b = nbl; b >= 1; b -= 2) {
    bwd_block4<KQ, NT>(L, fac, b, pa);
    bwd_block4<KQ, NT>(L, fac, b - 1, pb);
  }
  // (the last backward block ends with a barrier)
  double2 pq = write_outputs<NT>(a, m, L.v, w, uk);
  if (a.mode == MODE_ITER && a.pqpart) {
    __shared__ double2 shq[NT / WL];
    pq = wave_csum(pq);
    if ((t & (WL - 1)) == 0) shq[t / WL] = pq;
    __syncthreads();
    if (t == 0) {
      double2 sum = shq[0];
      for (int v = 1; v < NT / WL; v++) sum = cadd(sum, shq[v]);
      a.pqpart[w] = sum;
    }
  }
}

// (four waves hold two prefetch sets of KQ loads per thread: the 1,024-entry
// blocks need more than 128 VGPRs, so three workgroups per CU, the wide
// blocks two, which is what their LDS allows anyway)
template <int KQ, int NT>  // prefetch loads per thread; threads: row blocks of at most NT*KQ entries
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(NT == 256 ? (KQ == 8 ? 2 : KQ == 4 ? 3 : 4) : 3))) void k_solve4(SolveArgs a) {
  to_system(a);
  const bool check_done = a.mode == MODE_ITER && a.done;
  const int done0 = check_done ? *a.done : 0;
  int w = 0, uk = 0;
  if (!solve_item(a, w, uk)) return;
  const SubMeta m = a.meta[w];
  if (a.mode == MODE_UNIT && uk >= m.ng) return;
  extern __shared__ double2 lds[];
  solve4_body<KQ, NT>(a, m, w, uk, solve_lds(a, lds, NT * KQ, 2 * MINV), done0);
}

// ---------------------------------------------------------------------------
// Persistent form with size-bucketed queues: pgrid workgroups per system
// (as many as are resident at once) claim subdomains until none is left.
// NT = 256: first the large subdomains (porder[0, nlarge), largest first),
// each solved by the whole workgroup as in k_solve4 (KL loads per thread,
// blocks of at most 256 KL entries); then the small ones, each wave on its
// own subdomain in its own part of the workgroup's LDS as in k_solve (KS
// loads per lane), so a CU holds one large subdomain or four small ones in
// the same LDS.  NT = 64: small subdomains only.  Each subdomain's outputs
// go where it says, whichever workgroup solves it, and its sums are those of
// the kernel form that solves it.  The workgroup that ends last puts the
// queue back to zero for the next launch (no memset, so graphs replay it).
template <int KL, int KS, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 64 ? 2 : 1))) void
k_solveq(SolveArgs a) {
  int *q = a.queue + 4 * blockIdx.y;
  to_system(a);
  // a finished system: every workgroup of it stops before claiming
  // anything, so the queue stays zero
  if (a.mode == MODE_ITER && a.done && *a.done) return;
  extern __shared__ double2 lds[];
  const int t = threadIdx.x;
  if (NT > WL) {
    __shared__ int next;
    for (;;) {
      if (t == 0) next = atomicAdd(q, 1);
      __syncthreads();
      // (uniform: in scalar registers, so the subdomain's metadata is too)
      const int i = __builtin_amdgcn_readfirstlane(next);
      __syncthreads();  // every thread has read the claim before the next one is written
      if (i >= a.nlarge) break;
      const int w = a.porder[i];
      const SubMeta m = a.meta[w];
      solve4_body<KL, NT>(a, m, w, 0, solve_lds(lds, a.nmax_l, a.nblkmax_l, NT * KL, 2 * MINV), 0);
    }
    __syncthreads();  // the large items' LDS is free for the waves' own parts
  }
  const int lane = t & (WL - 1);
  double2 *mine = lds + (size_t)a.wave_lds * (t / WL);
  const int nsmall = a.nwork - a.nlarge;
  for (;;) {
    int i = 0;
    if (lane == 0) i = atomicAdd(q + 1, 1);
    i = __builtin_amdgcn_readfirstlane(i);  // lane 0's claim, uniform
    if (i >= nsmall) break;
    const int w = a.porder[a.nlarge + i];
    const SubMeta m = a.meta[w];
    solve1_body<KS, true>(a, m, w, 0, solve_lds(mine, a.nmax, a.nblkmax, WL * KS, MINV), 0);
    wave_sync();
  }
  __syncthreads();
  if (t == 0 && atomicAdd(q + 2, 1) == (int)gridDim.x - 1) {
    // every workgroup has made its last claims
    atomicExch(q, 0);
    atomicExch(q + 1, 0);
    atomicExch(q + 2, 0);
  }
}

// Inverse of every row block's unit lower triangle L_BB (the factor's strict
// lower entries inside the block, unit diagonal), packed by rows: entry
// (r, q), q < r, at r(r-1)/2 + q, the rest of the MINV slots zero.
// (L^-1)_rq = -(L_rq + sum_{q<j<r} L_rj (L^-1)_jq).  One workgroup per
// subdomain, one thread per block, every row-block partition (long, short,
// wide).
__global__ __launch_bounds__(64) void k_blockinv(const SubMeta *__restrict__ meta, const int *__restrict__ idx1,
                                                 const int *__restrict__ rowbeg, const int *__restrict__ blk_row,
                                                 const double2 *__restrict__ fac_all, double2 *__restrict__ minv) {
  const SubMeta m = meta[blockIdx.x];
  const double2 *fac = fac_all + m.fac0;
  for (int part = 0; part < 3; part++) {
    const int nblk = part == 0 ? m.nblk : part == 1 ? m.nblk_s : m.nblk_w;
    const int blk0 = part == 0 ? m.blk0 : part == 1 ? m.blk0_s : m.blk0_w;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
      const int i0 = blk_row[blk0 + b], nb = blk_row[blk0 + b + 1] - i0;
      double2 M[MINV];
      for (int t = 0; t < MINV; t++) M[t] = cz();
      for (int r = 1; r < nb; r++) {
        const int f = idx1[m.dof0 + i0 + r], rb = rowbeg[m.dof0 + i0 + r];
        for (int q = 0; q < r; q++) {
          double2 acc = (i0 + q >= f) ? fac[rb + (i0 + q - f)] : cz();
          for (int j = q + 1; j < r; j++)
            if (i0 + j >= f) acc = cfma(acc, fac[rb + (i0 + j - f)], M[(j * (j - 1)) / 2 + q]);
          M[(r * (r - 1)) / 2 + q] = csub(cz(), acc);
        }
      }
      double2 *out = minv + (long long)(blk0 + b) * MINV;
      for (int t = 0; t < MINV; t++) out[t] = M[t];
    }
  }
}

hipError_t launch_blockinv(const SubMeta *meta, int nsub, const int *idx1, const int *rowbeg, const int *blk_row,
                           const double2 *fac, double2 *minv, hipStream_t st) {
  if (nsub <= 0) return hipSuccess;
  k_blockinv<<<nsub, 64, 0, st>>>(meta, idx1, rowbeg, blk_row, fac, minv);
  return hipGetLastError();
}

size_t solve_lds_bytes(int nmax, int ngmax, int nblkmax, int pfk, int waves) {
  (void)ngmax;
  // the two- and four-wave forms: the same block capacity (64 pfk
  // entries), two slots for the inverse triangles, and the (p,q) partials
  // of their waves
  const size_t ints = sizeof(int) * ((size_t)2 * nmax + 1 + nblkmax + 1) + sizeof(BlkDesc) * ((size_t)nblkmax + 1);
  if (waves == 2 || waves == 4) return sizeof(double2) * ((size_t)nmax + (size_t)WL * pfk + 2 * MINV + waves) + ints;
  return sizeof(double2) * ((size_t)nmax + (size_t)WL * pfk + MINV) + ints;
}

size_t solve_persistent_lds(const SolveArgs &a) {
  // every wave's own part: wave_lds entries (at least its layout)
  const size_t one = std::max(solve_lds_bytes(a.nmax, a.ngmax, a.nblkmax, a.pfk, 1),
                              sizeof(double2) * (size_t)a.wave_lds);
  size_t b = one * (size_t)(a.pthreads / WL);
  if (a.pthreads > WL && sizeof(double2) * (size_t)a.wave_lds < solve_lds_bytes(a.nmax, a.ngmax, a.nblkmax, a.pfk, 1))
    return (size_t)1 << 30;  // a wave's part would overlap the next: never fits
  if (a.nlarge > 0) {
    // the large items' layout (its static (p,q) partials are not dynamic LDS)
    const size_t large = solve_lds_bytes(a.nmax_l, a.ngmax, a.nblkmax_l, a.pfk_l, 4) - sizeof(double2) * 4;
    b = std::max(b, large);
  }
  return b;
}

// the persistent kernel of the args' shape (null: not instantiated)
static const void *solveq_kernel(const SolveArgs &a) {
  // (the size-bucketed form, 256 threads with a large queue, is switched
  // off: an iterative solve through it went silent on the GPU and the cause
  // is not established; it is not instantiated)
  if (a.pthreads == 64 && a.nlarge == 0) {
    if (a.pfk == 8) return (const void *)k_solveq<4, 8, 64>;
    if (a.pfk == 16) return (const void *)k_solveq<4, 16, 64>;
  }
  return nullptr;
}

int solve_persistent_occupancy(const SolveArgs &a) {
  const void *k = solveq_kernel(a);
  const size_t lds = solve_persistent_lds(a);
  if (!k || lds > 160 * 1024) return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, a.pthreads, lds) != hipSuccess) return 0;
  return nb;
}

hipError_t launch_solve(const SolveArgs &a, int nwork, hipStream_t st) {
  if (nwork <= 0) return hipSuccess;
  if (a.pgrid > 0) {
    // (one subdomain per entry of porder; the caller's item count must match)
    const void *k = solveq_kernel(a);
    if (!k || !a.queue || !a.porder || a.mode == MODE_UNIT || nwork != a.nwork || a.nlarge < 0 ||
        a.nlarge > a.nwork || (a.pthreads == 64 && a.nlarge > 0))
      return hipErrorInvalidValue;
    const dim3 grid(a.pgrid, a.nsys > 0 ? a.nsys : 1);
    const size_t lds = solve_persistent_lds(a);
    if (a.pfk == 8)
      k_solveq<4, 8, 64><<<grid, 64, lds, st>>>(a);
    else
      k_solveq<4, 16, 64><<<grid, 64, lds, st>>>(a);
    return hipGetLastError();
  }
  const dim3 grid(nwork, a.nsys > 0 ? a.nsys : 1);
  if (a.waves == 2 || a.waves == 4) {
    // the static (p,q) partials are not part of the dynamic size
    const size_t lds = solve_lds_bytes(a.nmax, a.ngmax, a.nblkmax, a.pfk, a.waves) - sizeof(double2) * a.waves;
    if (a.waves == 4) {
      if (a.pfk == 8)
        k_solve4<2, 256><<<grid, 256, lds, st>>>(a);
      else if (a.pfk == 32)
        k_solve4<8, 256><<<grid, 256, lds, st>>>(a);  // the wide blocks (<= 2,048 entries)
      else
        k_solve4<4, 256><<<grid, 256, lds, st>>>(a);
    } else {
      if (a.pfk == 32) return hipErrorInvalidValue;  // wide blocks only with four waves
      if (a.pfk == 8)
        k_solve4<4, 128><<<grid, 128, lds, st>>>(a);
      else
        k_solve4<8, 128><<<grid, 128, lds, st>>>(a);
    }
    return hipGetLastError();
  }
  const size_t lds = solve_lds_bytes(a.nmax, a.ngmax, a.nblkmax, a.pfk, 1);
  if (a.pfk == 32) return hipErrorInvalidValue;
  if (a.pfk == 8)
    k_solve<8><<<grid, WL, lds, st>>>(a);
  else
    k_solve<16><<<grid, WL, lds, st>>>(a);
  return hipGetLastError();
}

}  // namespace fw
