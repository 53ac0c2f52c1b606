// fw_schur.hip - the interface matvec q = S p with explicit subdomain Schur
// complements.
//
// The Schur step of the reference (hddm_solver.h:559-622 with
// renewal_bc_sp, hddm_solver_sp.h:58-69, and COM_res_sp,
// hddm_solver_sp.h:28-54) applies, per subdomain and per interface
// iteration, y_G = (A x)_G with A_II x_I = -A_IG p_G, x_G = p_G; that is
// y = S_s p_s with the complex-symmetric Schur complement
// S_s = A_GG - A_GI A_II^-1 A_IG.  In explicit mode the S_s are formed once
// from the skyline factors (S_s = A_GG - W^T D^-1 W with W = L^-1 (-A_IG):
// the inverted 16-row diagonal blocks of L (k_blockinv16), blocked forward
// substitution of 16-column chunks of -A_IG (k_schur_fwd), then
// k_schur_gemm on the matrix cores and k_schur_agg; or the Schur step of
// k_solve on every interface unit
// vector, MODE_UNIT) and each iteration streams them once through k_smv: a
// batched, HBM-bound symmetric matvec instead of two dependent triangular
// sweeps per subdomain.
//
// Storage of S_s: the lower triangle in 16x16 tiles, row block by row block;
// tile (I,J), J <= I, starts at tile_off(I,J).  Inside a tile, entry (row r,
// column c) is in 8x8 sub-block k = (r>>3)*2 + (c>>3) at k*64 +
// ((r&7) | (c&7)<<3), so each of the four 64-lane loads of a tile reads 1 KB
// contiguously and lane l holds rows l&7, (l&7)+8 and columns l>>3, (l>>3)+8.
// Diagonal tiles hold the strict lower part and half the diagonal: their
// sub-block 1 (rows 0-7, columns 8-15) is entirely upper and not stored, and
// sub-blocks 0 and 3 are packed lower triangles, so they take 136 entries
// (lanes of an upper entry read zero).  When the last row block has 8 or
// fewer rows, its tiles keep only their top half (the diagonal one its packed
// sub-block 0): the rest would be padding (tile_at, smv_entries).  Every
// tile T is applied as T + T^T, which restores S_s exactly.
#include "fw_kernels.h"
#include "fw_device.h"

namespace fw {

namespace {

struct Ent {
  double x, y;
};
struct Tile {
  Ent e[4];  // e[a*2+b]: row (l&7)+8a, column (l>>3)+8b
};

// The tiles are streamed once per application and do not fit any cache
// level (hundreds of MB): non-temporal loads keep them from evicting the
// interface vectors and measured 12 % faster than ordinary loads.
// row block and column block of the t-th tile of the row-block-major sequence
__device__ __forceinline__ void tile_ij(long long t, int &I, int &J) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((long long)(i + 1) * (i + 2) / 2 <= t) i++;
  while ((long long)i * (i + 1) / 2 > t) i--;
  I = i;
  J = (int)(t - (long long)i * (i + 1) / 2);
}

// Tiles are loaded in sequence through a cursor (wave-uniform, scalar
// registers) that stops at the wave's last tile, so every load is
// unconditional.  A diagonal tile's missing sub-block 1 is read from the
// 64 zero entries at the head of the tile array (cache-resident: no HBM
// traffic, no select).
struct Cursor {
  long long t, tl;  // next tile, last tile of the wave
  int I, J;         // its row block and column block
  int nb;           // row blocks of the subdomain
  bool half;        // its last row block keeps the top half only
};
__device__ __forceinline__ void load_next(const double2 *__restrict__ base, const double2 *__restrict__ zero,
                                          Cursor &c, int lane, Tile &T) {
  const bool dg = c.I == c.J;
  const bool hf = c.half && c.I == c.nb - 1;  // bottom sub-blocks not stored: read zeros
  const double2 *t = base + tile_at(c.I, c.J, c.nb, c.half);
  const double2 *z = zero + lane;
  // diagonal tile: packed triangles (lane (r, c) = (lane&7, lane>>3), upper
  // lanes read zeros), full sub-block 2 at 36
  const int r8 = lane & 7, c8 = lane >> 3;
  const bool up = c8 > r8;
  const int tr = tri8(r8, c8);
  const double2 *q0 = dg ? (up ? z : t + tr) : t + lane;
  const double2 *q1 = dg ? z : t + 64 + lane;
  const double2 *q2 = hf ? z : (dg ? t + 36 + lane : t + 128 + lane);
  const double2 *q3 = (hf || (dg && up)) ? z : (dg ? t + 100 + tr : t + 192 + lane);
  T.e[0].x = __builtin_nontemporal_load(&q0->x);
  T.e[0].y = __builtin_nontemporal_load(&q0->y);
  T.e[1].x = __builtin_nontemporal_load(&q1->x);
  T.e[1].y = __builtin_nontemporal_load(&q1->y);
  T.e[2].x = __builtin_nontemporal_load(&q2->x);
  T.e[2].y = __builtin_nontemporal_load(&q2->y);
  T.e[3].x = __builtin_nontemporal_load(&q3->x);
  T.e[3].y = __builtin_nontemporal_load(&q3->y);
  if (c.t < c.tl) {
    c.t++;
    if (++c.J > c.I) {
      c.I++;
      c.J = 0;
    }
  }
}

__device__ __forceinline__ double2 to_c(const Ent &e) { return make_double2(e.x, e.y); }

// a double moved between lanes by a DPP pattern (two 32-bit moves)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over each aligned group of 8 lanes, in registers: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], then row_half_mirror (lane i <- lane 7-i of its
// 8-lane half-row).  Every lane of a group ends with the same sum.
__device__ __forceinline__ double sum8(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  return v;
}
__device__ __forceinline__ double2 csum8(double2 v) { return make_double2(sum8(v.x), sum8(v.y)); }

struct SmvState {
  int I, J;
  double2 q0, q1;    // row partials (rows r8, r8+8 of block I) over this lane's columns
  double2 pI0, pI1;  // p at rows r8, r8+8 of block I
};

// one tile: row contributions accumulate in registers until the row block
// ends; column contributions are summed over the 8 row lanes at once and
// added to the wave's private accumulator
__device__ __forceinline__ void tile_step(const Tile &T, SmvState &st, const double2 *pv, double2 *my, int nb,
                                          bool last, int lane) {
  const int r8 = lane & 7, c8 = lane >> 3;
  const double2 pJ0 = pv[st.J * 16 + c8], pJ1 = pv[st.J * 16 + c8 + 8];
  const double2 s00 = to_c(T.e[0]), s01 = to_c(T.e[1]), s10 = to_c(T.e[2]), s11 = to_c(T.e[3]);
  st.q0 = cfma(cfma(st.q0, s00, pJ0), s01, pJ1);
  st.q1 = cfma(cfma(st.q1, s10, pJ0), s11, pJ1);
  double2 u0 = cfma(cmul(s00, st.pI0), s10, st.pI1);  // column c8
  double2 u1 = cfma(cmul(s01, st.pI0), s11, st.pI1);  // column c8 + 8
  u0 = csum8(u0);
  u1 = csum8(u1);
  if (r8 == 0) {
    double2 *d = my + st.J * 16 + c8;
    d[0] = cadd(d[0], u0);
    d[8] = cadd(d[8], u1);
  }
  st.J++;
  if (st.J > st.I || last) {
    double2 q0 = st.q0, q1 = st.q1;
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) {
      q0.x += __shfl_xor(q0.x, o, 64);
      q0.y += __shfl_xor(q0.y, o, 64);
      q1.x += __shfl_xor(q1.x, o, 64);
      q1.y += __shfl_xor(q1.y, o, 64);
    }
    if (c8 == 0) {
      double2 *d = my + st.I * 16 + r8;
      d[0] = cadd(d[0], q0);
      d[8] = cadd(d[8], q1);
    }
    st.q0 = cz();
    st.q1 = cz();
    st.I++;
    st.J = 0;
    if (st.I < nb) {
      st.pI0 = pv[st.I * 16 + r8];
      st.pI1 = pv[st.I * 16 + r8 + 8];
    }
  }
}

}  // namespace

// the value arrays of system sys0 + blockIdx.y (index arrays are shared)
__device__ __forceinline__ void to_system(SmvArgs &a) {
  const long long s = a.sys0 + (long long)blockIdx.y;
  if (s == 0) return;
  a.smat += s * a.ss.smat;
  a.P += s * a.ss.g;
  if (a.Z) a.Z += s * a.ss.g;
  a.sc += s;
  if (a.done) a.done = &a.sc->done;
  if (a.part) a.part += s * 3LL * a.G;
  if (a.pqpart) a.pqpart += s * (long long)a.npq;
  a.yslot += s * a.ss.slot;
}

// The status of the interface iteration for the Schur matvec (false: the
// system is finished and the workgroup's work is discarded; beta: the
// coefficient of the next search direction).  When the previous iteration's
// status step is still pending (k_iter_ab ran, k_iter_c did not), every
// workgroup takes it from k_iter_ab's partials in k_iter_c's order, and
// workgroup 0 stages it for the next k_iter_ab to commit; no workgroup of
// this kernel reads the staged fields.  Contains barriers: call uniformly.
// With 256 threads each thread is its own virtual thread of that order, and
// its first four partials of each kind can be loaded before the workgroup's
// first tiles: vector loads complete in issue order, so waiting for them
// then does not wait for the tiles.
struct StatusPre {
  double2 rz[4];
  double mx[4];
};
__device__ __forceinline__ void status_prefetch(const SmvArgs &a, StatusPre &q) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int k = threadIdx.x + 256 * j;
    q.rz[j] = k < a.G ? a.part[a.G + k] : cz();
    q.mx[j] = k < a.G ? a.part[2 * a.G + k].x : 0.0;
  }
}
// parts_csum256 / parts_max256 of 256 threads with the first four entries
// of every thread already loaded (same order: t, t+256, ...)
__device__ __forceinline__ double2 csum256_pre(const double2 *__restrict__ p, int G, double2 *sh, const StatusPre &q) {
  const int t = threadIdx.x;
  double2 v = cz();
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (t + 256 * j < G) v = cadd(v, q.rz[j]);
  for (int k = t + 1024; k < G; k += 256) v = cadd(v, p[k]);
  v = wave_csum(v);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double2 r = cz();
#pragma unroll
  for (int k = 0; k < 4; k++) r = cadd(r, sh[k]);
  return r;
}
__device__ __forceinline__ double max256_pre(const double2 *__restrict__ p, int G, double *sh, const StatusPre &q) {
  const int t = threadIdx.x;
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (t + 256 * j < G) v = fmax(v, q.mx[j]);
  for (int k = t + 1024; k < G; k += 256) v = fmax(v, p[k].x);
  v = wave_max(v);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < 4; k++) r = fmax(r, sh[k]);
  return r;
}

__device__ __forceinline__ bool iter_status(const SmvArgs &a, int NT, double2 &beta, const StatusPre *pre) {
  Scal *s = a.sc;
  const int done = s->done, pending = s->pending, loop = s->loop, max_loop = s->max_loop;
  const double r0inv = s->r0inv, eps = s->eps;
  const double2 rzo = s->rz;
  beta = s->beta;
  // the partials are reduced whether or not the step is pending
  __shared__ double2 shz[4];
  __shared__ double shx[4];
  const double2 rzs = pre ? csum256_pre(a.part + a.G, a.G, shz, *pre) : parts_csum256(a.part + a.G, a.G, shz, NT);
  const double mxs = pre ? max256_pre(a.part + 2 * a.G, a.G, shx, *pre) : parts_max256(a.part + 2 * a.G, a.G, shx, NT);
  if (done) return false;
  if (!pending) return true;
  const StatusStep r = status_step(rzs, mxs, rzo, r0inv, eps, loop, max_loop);
  if (threadIdx.x == 0 && blockIdx.x == 0 && a.base == 0) {
    s->st_rz = rzs;
    s->st_beta = r.beta;
    s->st_resi = r.resi;
    s->st_conv = r.conv;
    s->st_done = r.done;
    s->st_fresh = 1;
  }
  beta = r.beta;
  return !r.done;
}

template <int W>  // waves per workgroup
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(5))) void k_smv(SmvArgs a) {
  constexpr int SMV_WAVES = W, NT = 64 * W;
  to_system(a);
  SmvWork wk;
  if (a.work) {
    wk = a.work[a.base + blockIdx.x];
  } else {
    const int w = a.order ? a.order[a.base + blockIdx.x] : a.base + blockIdx.x;
    wk.off = a.sm_off[w];
    wk.y0 = a.meta[w].g0;
    wk.gg = a.meta[w].g0;
    wk.ng = a.meta[w].ng;
    wk.pq = w;
    wk.t0 = 0;
    wk.t1 = 0;
  }
  const int ng = wk.ng, nb = (ng + 15) >> 4, npad = nb * 16, stride = a.nbmax * 16;
  extern __shared__ double2 lds[];
  double2 *pv = lds;          // [nbmax*16] interface values of the subdomain
  double2 *acc = pv + stride; // [SMV_WAVES][nbmax*16] per-wave accumulators
  // the wave index is uniform within a wave: readfirstlane keeps it, the
  // tile ranges and the load cursor in scalar registers
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;

  if (ng == 0) {
    double2 beta;
    if (a.done && a.Z && blockIdx.x == 0 && a.base == 0) (void)iter_status(a, NT, beta, nullptr);  // it may have to stage
    if (a.pqpart && tid == 0) a.pqpart[wk.pq] = cz();
    return;
  }

  // contiguous, equal shares of the item's tiles (row-block major; all of
  // the subdomain's, or the range [wk.t0, wk.t1) of a split one); the
  // first two tiles of every wave go in flight before the interface values
  // are gathered, so their latency overlaps the gather (a wave without tiles
  // loads tile 0, which every subdomain with an interface has)
  const long long ntile = (long long)nb * (nb + 1) / 2;
  const long long T0 = wk.t1 > wk.t0 ? wk.t0 : 0, T1 = wk.t1 > wk.t0 ? wk.t1 : ntile;
  const long long t0 = T0 + (T1 - T0) * wave / SMV_WAVES, t1 = T0 + (T1 - T0) * (wave + 1) / SMV_WAVES;
  const bool active = t0 < t1;
  const long long tl = active ? t1 - 1 : 0;
  const double2 *base = a.smat + wk.off;
  Cursor cur;
  cur.t = active ? t0 : 0;
  cur.tl = tl;
  cur.nb = nb;
  cur.half = smv_half(ng);
  tile_ij(cur.t, cur.I, cur.J);
  const double2 *zero = a.smat;  // the 64 zero entries ahead of every subdomain's tiles
  // requested before the tiles (loads complete in issue order): the status
  // partials of an interface iteration (256 threads) and the ids of this
  // thread's first two slots
  StatusPre sp;
  const bool pre = W == 4 && a.done && a.Z;
  if (pre) status_prefetch(a, sp);
  const int k0 = tid, k1 = tid + NT;
  const int gid0 = k0 < ng ? a.g_gid[wk.gg + k0] : 0;
  const int gid1 = k1 < ng ? a.g_gid[wk.gg + k1] : 0;
  Tile A, B;
  load_next(base, zero, cur, lane, A);
  load_next(base, zero, cur, lane, B);
  // the values of those two slots (they do not depend on the status) follow
  // the first tiles, ahead of the status fields
  double2 pa = cz(), za = cz(), pb = cz(), zb = cz();
  if (k0 < ng) {
    pa = a.P[gid0];
    if (a.Z) za = a.Z[gid0];
  }
  if (k1 < ng) {
    pb = a.P[gid1];
    if (a.Z) zb = a.Z[gid1];
  }
  // the status is read once the first tiles are in flight (a finished
  // system's workgroups discard them)
  double2 beta = cz();
  if (a.done) {
    if (a.Z) {
      if (!iter_status(a, NT, beta, pre ? &sp : nullptr)) return;
    } else if (*a.done) {
      return;
    }
  } else if (a.Z) {
    beta = a.sc->beta;  // a timing run: the current status as it stands
  }

  // interface values: P, or the next search direction z + beta p while the
  // interface iteration runs (formed here, not stored, see k_iter_ab)
  for (int k = tid; k < npad; k += NT) {
    double2 p = cz();
    if (k < ng) {
      double2 z;
      if (k == k0) {
        p = pa;
        z = za;
      } else if (k == k1) {
        p = pb;
        z = zb;
      } else {
        const int gid = a.g_gid[wk.gg + k];
        p = a.P[gid];
        z = a.Z ? a.Z[gid] : cz();
      }
      if (a.Z) p = zbp(z, beta, p);
    }
    pv[k] = p;
  }
  for (int k = tid; k < SMV_WAVES * stride; k += NT) acc[k] = cz();
  __syncthreads();

  if (active) {
    double2 *my = acc + wave * stride;
    SmvState st;
    tile_ij(t0, st.I, st.J);
    st.q0 = cz();
    st.q1 = cz();
    st.pI0 = pv[st.I * 16 + (lane & 7)];
    st.pI1 = pv[st.I * 16 + (lane & 7) + 8];
    // two tiles in flight in two statically named register sets; the loads
    // are unconditional (clamped to the wave's last tile) so the compiler can
    // wait for one set while the other stays in flight
    for (long long t = t0; t < t1; t += 2) {
      tile_step(A, st, pv, my, nb, t == tl, lane);
      load_next(base, zero, cur, lane, A);
      if (t + 1 < t1) tile_step(B, st, pv, my, nb, t + 1 == tl, lane);
      load_next(base, zero, cur, lane, B);
    }
  }
  __syncthreads();
  // y, and in an explicit iteration this subdomain's part of (p,q) =
  // sum over its slots of p y (the interface sum of p q, regrouped by
  // subdomain; fixed order: per thread ascending, waves in order)
  double2 pq = cz();
  for (int k = tid; k < ng; k += NT) {
    double2 y = cadd(acc[k], acc[stride + k]);
    for (int v = 2; v < SMV_WAVES; v++) y = cadd(y, acc[v * stride + k]);
    a.yslot[wk.y0 + k] = y;
    pq = cadd(pq, cmul(pv[k], y));
  }
  if (a.pqpart) {
    __shared__ double2 shq[SMV_WAVES];
    pq = wave_csum(pq);
    if (lane == 0) shq[wave] = pq;
    __syncthreads();
    if (tid == 0) {
      double2 t = shq[0];
      for (int v = 1; v < SMV_WAVES; v++) t = cadd(t, shq[v]);
      a.pqpart[wk.pq] = t;
    }
  }
}

// the status of an interface iteration as the first work item of a
// workgroup needs it (false: the system is finished; beta: the coefficient
// of the next search direction).  Contains barriers (iter_status): call
// uniformly.
__device__ __forceinline__ bool smv_status(const SmvArgs &a, int NT, double2 &beta, const StatusPre *pre) {
  if (a.done) {
    if (a.Z) return iter_status(a, NT, beta, pre);
    return *a.done == 0;
  }
  if (a.Z) beta = a.sc->beta;  // a timing run: the current status as it stands
  return true;
}

// work item i of the launch: an entry of the work list, or a subdomain of
// the order
__device__ __forceinline__ SmvWork smv_item(const SmvArgs &a, int i) {
  SmvWork wk;
  if (a.work) {
    wk = a.work[a.base + i];
  } else {
    const int w = a.order ? a.order[a.base + i] : a.base + i;
    wk.off = a.sm_off[w];
    wk.y0 = a.meta[w].g0;
    wk.gg = a.meta[w].g0;
    wk.ng = a.meta[w].ng;
    wk.pq = w;
    wk.t0 = 0;
    wk.t1 = 0;
  }
  return wk;
}

// One work item of the persistent matvec (k_smvq): y = S_s p over the
// item's tiles, as k_smv computes it.  first: the workgroup's first item,
// which also takes the status of the iteration (after its first tiles are
// in flight, or at once when the item has no interface DOFs) into beta;
// later items use that beta.  Returns false when the system is finished
// (the workgroup stops).  Contains barriers: call uniformly.  (k_smv keeps
// its own copy of this body: at five waves per SIMD it sits at the register
// limit, and the shared form spilled there.)
template <int W>
__device__ __forceinline__ bool smv_run(const SmvArgs &a, const SmvWork &wk, bool first, double2 &beta) {
  constexpr int SMV_WAVES = W, NT = 64 * W;
  const int ng = wk.ng, nb = (ng + 15) >> 4, npad = nb * 16, stride = a.nbmax * 16;
  extern __shared__ double2 lds[];
  double2 *pv = lds;          // [nbmax*16] interface values of the subdomain
  double2 *acc = pv + stride; // [SMV_WAVES][nbmax*16] per-wave accumulators
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;

  if (ng == 0) {
    if (first && !smv_status(a, NT, beta, nullptr)) return false;
    if (a.pqpart && tid == 0) a.pqpart[wk.pq] = cz();
    return true;
  }

  // contiguous, equal shares of the item's tiles (row-block major; all of
  // the subdomain's, or the range [wk.t0, wk.t1) of a split one); the
  // first two tiles of every wave go in flight before the interface values
  // are gathered, so their latency overlaps the gather (a wave without tiles
  // loads tile 0, which every subdomain with an interface has)
  const long long ntile = (long long)nb * (nb + 1) / 2;
  const long long T0 = wk.t1 > wk.t0 ? wk.t0 : 0, T1 = wk.t1 > wk.t0 ? wk.t1 : ntile;
  const long long t0 = T0 + (T1 - T0) * wave / SMV_WAVES, t1 = T0 + (T1 - T0) * (wave + 1) / SMV_WAVES;
  const bool active = t0 < t1;
  const long long tl = active ? t1 - 1 : 0;
  const double2 *base = a.smat + wk.off;
  Cursor cur;
  cur.t = active ? t0 : 0;
  cur.tl = tl;
  cur.nb = nb;
  cur.half = smv_half(ng);
  tile_ij(cur.t, cur.I, cur.J);
  const double2 *zero = a.smat;  // the 64 zero entries ahead of every subdomain's tiles
  // requested before the tiles (loads complete in issue order): the status
  // partials of an interface iteration (256 threads, first item) and the ids
  // of this thread's first two slots
  StatusPre sp;
  const bool pre = W == 4 && first && a.done && a.Z;
  if (pre) status_prefetch(a, sp);
  const int k0 = tid, k1 = tid + NT;
  const int gid0 = k0 < ng ? a.g_gid[wk.gg + k0] : 0;
  const int gid1 = k1 < ng ? a.g_gid[wk.gg + k1] : 0;
  Tile A, B;
  load_next(base, zero, cur, lane, A);
  load_next(base, zero, cur, lane, B);
  // the values of those two slots (they do not depend on the status) follow
  // the first tiles, ahead of the status fields
  double2 pa = cz(), za = cz(), pb = cz(), zb = cz();
  if (k0 < ng) {
    pa = a.P[gid0];
    if (a.Z) za = a.Z[gid0];
  }
  if (k1 < ng) {
    pb = a.P[gid1];
    if (a.Z) zb = a.Z[gid1];
  }
  // the status is read once the first tiles are in flight (a finished
  // system's workgroups discard them)
  // (two calls, each with a fixed pointer, so the prefetched partials stay
  // in registers)
  if (first && !(pre ? smv_status(a, NT, beta, &sp) : smv_status(a, NT, beta, nullptr))) return false;

  // interface values: P, or the next search direction z + beta p while the
  // interface iteration runs (formed here, not stored, see k_iter_ab)
  for (int k = tid; k < npad; k += NT) {
    double2 p = cz();
    if (k < ng) {
      double2 z;
      if (k == k0) {
        p = pa;
        z = za;
      } else if (k == k1) {
        p = pb;
        z = zb;
      } else {
        const int gid = a.g_gid[wk.gg + k];
        p = a.P[gid];
        z = a.Z ? a.Z[gid] : cz();
      }
      if (a.Z) p = zbp(z, beta, p);
    }
    pv[k] = p;
  }
  for (int k = tid; k < SMV_WAVES * stride; k += NT) acc[k] = cz();
  __syncthreads();

  if (active) {
    double2 *my = acc + wave * stride;
    SmvState st;
    tile_ij(t0, st.I, st.J);
    st.q0 = cz();
    st.q1 = cz();
    st.pI0 = pv[st.I * 16 + (lane & 7)];
    st.pI1 = pv[st.I * 16 + (lane & 7) + 8];
    // two tiles in flight in two statically named register sets; the loads
    // are unconditional (clamped to the wave's last tile) so the compiler can
    // wait for one set while the other stays in flight
    for (long long t = t0; t < t1; t += 2) {
      tile_step(A, st, pv, my, nb, t == tl, lane);
      load_next(base, zero, cur, lane, A);
      if (t + 1 < t1) tile_step(B, st, pv, my, nb, t + 1 == tl, lane);
      load_next(base, zero, cur, lane, B);
    }
  }
  __syncthreads();
  // y, and in an explicit iteration this subdomain's part of (p,q) =
  // sum over its slots of p y (the interface sum of p q, regrouped by
  // subdomain; fixed order: per thread ascending, waves in order)
  double2 pq = cz();
  for (int k = tid; k < ng; k += NT) {
    double2 y = cadd(acc[k], acc[stride + k]);
    for (int v = 2; v < SMV_WAVES; v++) y = cadd(y, acc[v * stride + k]);
    a.yslot[wk.y0 + k] = y;
    pq = cadd(pq, cmul(pv[k], y));
  }
  if (a.pqpart) {
    __shared__ double2 shq[SMV_WAVES];
    pq = wave_csum(pq);
    if (lane == 0) shq[wave] = pq;
    __syncthreads();
    if (tid == 0) {
      double2 t = shq[0];
      for (int v = 1; v < SMV_WAVES; v++) t = cadd(t, shq[v]);
      a.pqpart[wk.pq] = t;
    }
  }
  return true;
}

// Persistent form: pgrid workgroups per system, workgroup b takes item b,
// then claims the next unclaimed item from the system's queue (items in
// work order, largest first, so the claims are a longest-first list
// schedule) until none is left.  Each item's outputs go where the item
// says, whichever workgroup computes it, so the results are those of k_smv.
// The status of the iteration is taken once per workgroup (with its first
// item) instead of once per item.  The workgroup that ends last puts the
// queue back to zero for the next launch (no memset, so graphs replay it).
template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void k_smvq(SmvArgs a, int nwork) {
  int *q = a.queue + 2 * blockIdx.y;
  to_system(a);
  __shared__ int next;
  double2 beta = cz();
  // a finished system: every workgroup of it reads the same status at its
  // first item and stops before claiming anything, so the queue stays zero
  if (!smv_run<W>(a, smv_item(a, blockIdx.x), true, beta)) return;
  for (;;) {
    __syncthreads();  // the item's LDS reads are done, and the last claim was read
    if (threadIdx.x == 0) next = (int)gridDim.x + atomicAdd(q, 1);
    __syncthreads();
    const int i = next;
    if (i >= nwork) break;
    (void)smv_run<W>(a, smv_item(a, i), false, beta);
  }
  if (threadIdx.x == 0 && atomicAdd(q + 1, 1) == (int)gridDim.x - 1) {
    // every workgroup has made its last claim
    atomicExch(q, 0);
    atomicExch(q + 1, 0);
  }
}

// ---------------------------------------------------------------------------
// GEMM formation of S_s.  With the Dirichlet matrix factored as L D L^T and
// W = L^-1 (-A_IG), S_s = A_GG - W^T D^-1 W.  The interface and essential
// rows of W are zero (their rows of L are trivial, their loads zero), so W is
// formed over the rows up to the last interior one (nsw) only, in chunks of
// 16 interface columns (row i of chunk c at (c nsw + i) 16, the 16 columns of
// a row 256 contiguous bytes; columns past ng are zero).
// All three steps run on the FP64 matrix instruction v_mfma_f64_16x16x4f64:
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane
// j + 16 (i mod 4), register i / 4.  Register v of a result therefore holds
// rows 4v..4v+3, which is the B operand of the k-step over those rows: a
// block just solved feeds the next block's products from registers.
// A complex product takes four real ones: re += Ar Br - Ai Bi,
// im += Ar Bi + Ai Br.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma64(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// L_BB^-1 of every uniform 16-row block of the swept rows (unit lower
// triangle; rows past nsw are identity rows), 16x16 row-major:
// (L^-1)_rq = -(L_rq + sum_{q<j<r} L_rj (L^-1)_jq), one lane per column q,
// four blocks per wavefront.
__global__ __launch_bounds__(64) void k_blockinv16(BlockInv16Args a) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 4);
  if (g >= a.nblk) return;
  const int2 sb = a.blist[g];
  const SubMeta m = a.meta[sb.x];
  const int i0 = 16 * sb.y, nb = min(16, a.nsw[sb.x] - i0);
  const int q = threadIdx.x & 15;
  const int *idx1 = a.idx1 + m.dof0, *rowbeg = a.rowbeg + m.dof0;
  const double2 *fac = a.fac + m.fac0;
  double2 x[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const bool rv = r < nb;
    const int f = rv ? idx1[i0 + r] : 0, rb = rv ? rowbeg[i0 + r] : 0;
    double2 acc = cz();
    if (r > q) {
      if (rv && i0 + q >= f) acc = fac[rb + (i0 + q - f)];
#pragma unroll
      for (int j = 1; j < r; j++)
        if (j > q && rv && i0 + j >= f) acc = cfma(acc, fac[rb + (i0 + j - f)], x[j]);
    }
    x[r] = r < q ? cz() : (r == q ? make_double2(1.0, 0.0) : csub(cz(), acc));
  }
  double2 *out = a.m16 + (long long)g * 256;
#pragma unroll
  for (int r = 0; r < 16; r++) out[r * 16 + q] = x[r];
}

// W = L^-1 (-A_IG), 16 columns per wavefront, by blocked forward
// substitution over uniform 16-row blocks: X = B_R - L[R, <R] W[<R] (the
// off-block products, k-steps of 4 columns: the previous block's columns
// from registers, older ones from this chunk of W as stored), W_R = M_R X.
// B_R is the unit-vector load of the Schur step (hddm_solver.h:559-622 with
// renewal_bc_sp): for an interior row, minus its coupling to the column's
// interface DOF.  Rows of W are read only after being stored by the same
// wavefront one block earlier or more; its vector memory operations
// complete in issue order, and the row loads wait on loads issued after
// those stores.
__global__ __launch_bounds__(64) void k_schur_fwd(FwdArgs a) {
  const int2 pr = a.work[blockIdx.x];
  if (pr.x < 0) return;
  const int s = pr.x, ch = pr.y;
  const SubMeta m = a.meta[s];
  const int nsw = a.nsw[s];
  const int l = threadIdx.x, li = l & 15, lk = l >> 4;
  const int col = 16 * ch + li;  // interface column of this lane's B and D entries
  const int *idx1 = a.idx1 + m.dof0, *rowbeg = a.rowbeg + m.dof0;
  const double2 *fac = a.fac + m.fac0;
  double2 *W = a.wmat + a.w_off[s] + (long long)ch * nsw * 16;  // row i at W + 16 i
  const double2 *M = a.m16 + a.m16_off[s] * 256;
  mfma_d4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};  // the previous block's rows of W
  for (int r0 = 0; r0 < nsw; r0 += 16) {
    const int nb = min(16, nsw - r0);
    // this lane's row as the A operand: row li of the block
    const bool rv = li < nb;
    const int row = r0 + (rv ? li : 0);
    const int f = rv ? idx1[row] : r0, rb = rowbeg[row];
    int cmin = f;  // first column any row of the block reaches
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) cmin = min(cmin, __shfl_xor(cmin, o, 64));
    // the load rows (D layout: register v holds row 4v + lk, column li)
    mfma_d4 xr, xi;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = 4 * v + lk;
      double2 b = cz();
      if (i < nb) {
        const int d = m.dof0 + r0 + i;
        if (a.bcflag[d] == 0)
          for (int c = a.cp_ptr[d]; c < a.cp_ptr[d + 1]; c++)
            if (a.cp_k[c] == col) b = csub(b, a.cp_val[c]);
      }
      xr[v] = b.x;
      xi[v] = b.y;
    }
    // off-block products L[R, <R] W[<R] in four accumulators
    mfma_d4 arr = {0.0, 0.0, 0.0, 0.0}, aii = arr, ari = arr, air = arr;
    if (r0 > 0) {
      const int p0 = r0 - 16;  // the previous block (always 16 rows)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = p0 + 4 * v + lk;
        const double2 L = (rv && c >= f) ? fac[rb + (c - f)] : cz();
        arr = mfma64(L.x, wr[v], arr);
        aii = mfma64(L.y, wi[v], aii);
        ari = mfma64(L.x, wi[v], ari);
        air = mfma64(L.y, wr[v], air);
      }
      for (int c0 = p0 - 4; c0 + 3 >= cmin; c0 -= 4) {
        const int c = c0 + lk;
        const bool cv = c >= 0;
        const double2 L = (rv && cv && c >= f) ? fac[rb + (c - f)] : cz();
        const double2 w = cv ? W[16 * c + li] : cz();
        arr = mfma64(L.x, w.x, arr);
        aii = mfma64(L.y, w.y, aii);
        ari = mfma64(L.x, w.y, ari);
        air = mfma64(L.y, w.x, air);
      }
    }
    xr = xr - arr + aii;
    xi = xi - ari - air;
    // W_R = M_R X
    const double2 *Mb = M + (r0 >> 4) * 256 + li * 16 + lk;
    mfma_d4 orr = {0.0, 0.0, 0.0, 0.0}, oii = orr, ori = orr, oir = orr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const double2 mm = Mb[4 * v];
      orr = mfma64(mm.x, xr[v], orr);
      oii = mfma64(mm.y, xi[v], oii);
      ori = mfma64(mm.x, xi[v], ori);
      oir = mfma64(mm.y, xr[v], oir);
    }
    wr = orr - oii;
    wi = ori + oir;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = 4 * v + lk;
      if (i < nb) W[16 * (r0 + i) + li] = make_double2(wr[v], wi[v]);
    }
  }
}

// S_s tiles: one wavefront per 16x16 lower tile (I, J) accumulates
// C = W_I^T D^-1 W_J over the swept rows, D^-1 applied to the B operand as
// it is loaded; the tile then holds -C (diagonal halved) and k_schur_agg
// adds A_GG.
__global__ __launch_bounds__(256) void k_schur_gemm(GemmArgs a) {
  const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid >= a.ntile) return;
  const int4 t = a.tiles[wid];
  if (t.x < 0) return;
  const SubMeta m = a.meta[t.x];
  const int ng = m.ng, nsw = a.nsw[t.x], I = t.y, J = t.z;
  const double2 *W = a.wmat + a.w_off[t.x];
  const double2 *dinv = a.dinv + m.dof0;
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
  const double2 *wi = W + (long long)I * nsw * 16 + li;  // row k of chunk I at wi[16 k]
  const double2 *wj = W + (long long)J * nsw * 16 + li;
  mfma_d4 cre = {0.0, 0.0, 0.0, 0.0}, cim = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < nsw; k0 += 4) {
    const int k = k0 + lk;
    const bool in = k < nsw;
    const int kc = in ? k : 0;
    const double2 x = wi[16 * kc], y = wj[16 * kc], d = dinv[kc];
    const double ar = in ? x.x : 0.0, ai = in ? x.y : 0.0;
    const double2 v = cmul(d, y);
    const double br = in ? v.x : 0.0, bi = in ? v.y : 0.0;
    cre = mfma64(ar, br, cre);
    cre = mfma64(-ai, bi, cre);
    cim = mfma64(ar, bi, cim);
    cim = mfma64(ai, br, cim);
  }
  const int nb = (ng + 15) >> 4;
  const bool half = smv_half(ng);
  double2 *dst = a.smat + a.sm_off[t.x];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int ri = 4 * v + lk, c = li;  // C[ri][c] of this tile
    const int gi = 16 * I + ri, gj = 16 * J + c;
    if (gi < ng && gj < ng && (I > J || ri >= c)) {
      const double sc = gi == gj ? -0.5 : -1.0;
      dst[tile_pos(I, J, ri, c, nb, half)] = make_double2(sc * cre[v], sc * cim[v]);
    }
  }
}

// A_GG into the tiles: row i's interface-row entries at interface columns
// j <= i (the diagonal halved), added to -W^T D^-1 W; one workgroup per
// subdomain, one thread per interface row (each tile entry has one writer)
__global__ __launch_bounds__(256) void k_schur_agg(AggArgs a) {
  const int w = a.subs ? a.subs[blockIdx.x] : (int)blockIdx.x;
  const SubMeta m = a.meta[w];
  const int ng = m.ng, nb = (ng + 15) >> 4;
  const bool half = smv_half(ng);
  double2 *dst = a.smat + a.sm_off[w];
  for (int i = threadIdx.x; i < ng; i += blockDim.x) {
    const int sl = m.g0 + i;
    for (int c = a.gr_ptr[sl]; c < a.gr_ptr[sl + 1]; c++) {
      const int d = m.dof0 + a.gr_col[c];
      if (a.bcflag[d] != 1) continue;
      const int j = a.kidx[d];
      if (j > i) continue;
      double2 g = a.gr_val[c];
      if (j == i) g = cscale(g, 0.5);
      double2 *e = dst + tile_pos(i >> 4, j >> 4, i & 15, j & 15, nb, half);
      *e = cadd(*e, g);
    }
  }
}

hipError_t launch_blockinv16(const BlockInv16Args &a, hipStream_t st) {
  if (a.nblk <= 0) return hipSuccess;
  k_blockinv16<<<(a.nblk + 3) / 4, 64, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_schur_fwd(const FwdArgs &a, hipStream_t st) {
  if (a.nwork <= 0) return hipSuccess;
  k_schur_fwd<<<a.nwork, 64, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_schur_gemm(const GemmArgs &a, hipStream_t st) {
  if (a.ntile <= 0) return hipSuccess;
  k_schur_gemm<<<(a.ntile + 3) / 4, 256, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_schur_agg(const AggArgs &a, int nsub, hipStream_t st) {
  if (nsub <= 0) return hipSuccess;
  k_schur_agg<<<nsub, 256, 0, st>>>(a);
  return hipGetLastError();
}

size_t smv_lds_bytes(int nbmax, int waves) { return sizeof(double2) * (size_t)(1 + waves) * nbmax * 16; }

int smv_persistent_grid(int waves, int nbmax, int ncu) {
  const size_t lds = smv_lds_bytes(nbmax, waves);
  int nb = 0;
  hipError_t e;
  switch (waves) {
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_smvq<2>, 128, lds); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_smvq<8>, 512, lds); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_smvq<4>, 256, lds); break;
  }
  return (e == hipSuccess && nb > 0) ? nb * ncu : 0;
}

hipError_t launch_smv(const SmvArgs &a, int nwork, hipStream_t st) {
  if (nwork <= 0) return hipSuccess;
  const size_t lds = smv_lds_bytes(a.nbmax, a.waves);
  if (a.pgrid > 0) {
    if (!a.queue || a.base != 0) return hipErrorInvalidValue;
    const dim3 grid(std::min(nwork, a.pgrid), a.nsys > 0 ? a.nsys : 1);
    switch (a.waves) {
      case 2: k_smvq<2><<<grid, 128, lds, st>>>(a, nwork); break;
      case 8: k_smvq<8><<<grid, 512, lds, st>>>(a, nwork); break;
      default: k_smvq<4><<<grid, 256, lds, st>>>(a, nwork); break;
    }
    return hipGetLastError();
  }
  const dim3 grid(nwork, a.nsys > 0 ? a.nsys : 1);
  switch (a.waves) {
    case 2: k_smv<2><<<grid, 128, lds, st>>>(a); break;
    case 8: k_smv<8><<<grid, 512, lds, st>>>(a); break;
    default: k_smv<4><<<grid, 256, lds, st>>>(a); break;
  }
  return hipGetLastError();
}

}  // namespace fw
