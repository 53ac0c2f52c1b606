// fw_kernels.h - kernel argument blocks and launchers of libfwhip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fw {

struct Scal;

// per-subdomain metadata, one record per subdomain
struct SubMeta {
  int n;          // edge DOFs
  int ng;         // interface DOFs (slots)
  int nblk;       // row blocks of the skyline sweeps
  int dof0;       // offset into per-DOF arrays
  int g0;         // offset into per-slot arrays
  int blk0;       // offset into blk_row (nblk+1 entries)
  long long fac0; // offset into the skyline array (complex entries)
  long long nlen; // skyline length
  int nnz0;       // offset into the lower-triangle CSR arrays (entries)
  int nnz;        // entries
  int nblk_s;     // short row blocks (<= 64*8 entries; 0 when a row is wider)
  int blk0_s;     // offset of the short blocks in blk_row
  int nblk_w;     // wide row blocks (<= 64*32 entries, for four waves; 0 when not built)
  int blk0_w;     // offset of the wide blocks in blk_row
};

// MODE_UNIT: the Schur step applied to interface unit vector unit_k, output
// written into the explicit Schur-complement tiles (fw_schur.hip)
enum SolveMode { MODE_FIRST = 0, MODE_ITER = 1, MODE_LAST = 2, MODE_RHS = 3, MODE_UNIT = 4 };

// explicit Schur complements: 16x16 tiles of the lower triangle, row block
// by row block; off-diagonal tiles 256 entries (four 8x8 sub-blocks),
// diagonal tiles 136: their two diagonal 8x8 sub-blocks as packed lower
// triangles (36 each, entry (r, c), c <= r, at r(r+1)/2 + c) around the full
// sub-block 2 (rows 8-15, columns 0-7, at 36); sub-block 1 is upper
constexpr int SMV_TILE = 256;
constexpr int SMV_DIAG = 136;
__host__ __device__ __forceinline__ long long tile_off(int I, int J) {
  return 128LL * I * (I - 1) + (long long)SMV_DIAG * I + 256LL * J;
}
__host__ __device__ __forceinline__ int tri8(int r, int c) { return (r * (r + 1)) / 2 + c; }
// When the last row block holds 8 or fewer rows (ng mod 16 in 1..8), its
// tiles keep only their top half: sub-blocks 0 and 1 (128 entries), the
// diagonal tile its packed sub-block 0 (36 entries); the rest would be padding.
__host__ __device__ __forceinline__ bool smv_half(int ng) { return ng > 0 && ((ng - 1) & 15) < 8; }
__host__ __device__ __forceinline__ long long tile_at(int I, int J, int nb, bool half) {
  return (half && I == nb - 1) ? tile_off(I, 0) + 128LL * J : tile_off(I, J);
}
// position of S entry (row 16 I + ri, column 16 J + cj), J <= I, in the tiles
// of a subdomain with nb row blocks: off-diagonal tiles sub-block
// (ri>>3, cj>>3) of 64 entries; diagonal tiles packed triangles 0 (at 0) and
// 3 (at 100) around sub-block 2 (at 36); a half last row block has rows 0-7
__host__ __device__ __forceinline__ long long tile_pos(int I, int J, int ri, int cj, int nb, bool half) {
  const long long t = tile_at(I, J, nb, half);
  if (I != J) return t + ((ri >> 3) * 2 + (cj >> 3)) * 64 + ((ri & 7) | ((cj & 7) << 3));
  if (ri < 8) return t + tri8(ri, cj);
  if (cj < 8) return t + 36 + ((ri & 7) | (cj << 3));
  return t + 100 + tri8(ri - 8, cj - 8);
}
// tile entries of a subdomain with ng interface DOFs
__host__ __device__ __forceinline__ long long smv_entries(int ng) {
  const int nb = (ng + 15) >> 4;
  return smv_half(ng) ? tile_off(nb - 1, 0) + 128LL * (nb - 1) + 36 : tile_off(nb, 0);
}

// solve kernel limits: rows per block, prefetch registers per thread
constexpr int SOLVE_THREADS = 64;
constexpr int BMAX = 16;
constexpr int PFK = 16;                   // most prefetch loads per lane
constexpr int CAP = SOLVE_THREADS * PFK;  // most skyline entries per row block
constexpr int VT_HINT = 256;              // threads per block of the vector kernels
constexpr int MINV = 128;                 // entries per row block of the inverted diagonal
                                          // triangles (16*15/2 = 120 used, packed by rows)

// Systems of a batch (one per frequency of a sweep, say) share the mesh, the
// decomposition and every index array; their values are concatenated
// system-major with these strides (in entries).  A launch serves systems
// sys0 .. sys0 + nsys - 1, one per grid row (blockIdx.y).
struct SysStride {
  long long dof;   // per-DOF arrays: f, dinv, X
  long long nlen;  // skyline factor
  long long cpl;   // interior-interface coupling values
  long long grow;  // interface-row values
  long long slot;  // per-slot outputs yslot (with the matvec's slot copies)
  long long pcs;   // per-slot diagonal entries pcslot (the load's slots)
  long long g;     // interface vectors
  long long smat;  // Schur-complement tiles
  long long hist;  // residual history
  long long minv;  // inverted diagonal triangles of the row blocks
};

struct SolveArgs {
  int sys0, nsys;        // systems of this launch (grid rows)
  SysStride ss;
  const SubMeta *meta;
  const int *order;      // work order (may be null: identity)
  int base;              // first work item of this launch
  int mode;
  const int *done;       // iteration mode: skip when *done != 0 (may be null)
  int nmax, ngmax;       // LDS sizing
  int nblkmax;
  int pfk;               // prefetch loads per lane (8 or 16; 32 with four waves): blocks of <= 64*pfk entries
  int waves;             // waves per subdomain: 1, or 4 (k_solve4, the same blocks)
  const int *idx1;       // per DOF, local first column of the profile
  const int *rowbeg;     // per DOF, start of row i within the subdomain's skyline
  const double2 *fac;
  const double2 *dinv;   // per DOF, 1/d_i
  const int8_t *bcflag;  // per DOF: 0 interior, 1 interface, 2 essential
  const int *kidx;       // per DOF: interface slot (local) or -1
  const double2 *f;      // per DOF load
  const int *cp_ptr;     // per DOF (n_dof+1): interior-interface couplings
  const int *cp_k;
  const double2 *cp_val;
  const int *gr_ptr;     // per slot (n_slots+1): interface rows of A
  const int *gr_col;
  const double2 *gr_val;
  const int *g_nfid;     // per slot: local DOF
  const int *g_gid;      // per slot: global interface index
  const int *blk_row;    // per subdomain nblk+1 block start rows
  const double2 *minv;   // per block (index in blk_row) its inverted diagonal triangle
  const double2 *P;      // global interface values
  const double2 *Z;      // MODE_ITER: p = Z + beta P (null: p = P)
  const Scal *sc;        // beta of the interface iteration
  double2 *yslot;        // per slot output
  double2 *X;            // per DOF solution (MODE_LAST, MODE_RHS in/out)
  int unit_k;            // MODE_UNIT: local interface index of the unit vector
  const int2 *unit_list; // MODE_UNIT, if set: (subdomain, unit vector) per work item
  double2 *smat;         // MODE_UNIT: Schur-complement tiles
  const long long *sm_off;  // per subdomain first entry of its tiles
  double2 *pqpart;       // MODE_ITER of an interface iteration: per subdomain its (p,q) part (may be null)
  int npq;               // per system: subdomains
  // persistent form (pgrid > 0, batched launches other than MODE_UNIT):
  // pgrid workgroups of pthreads threads per system claim the subdomains
  // of porder from queue[4 s], queue[4 s + 1] (workgroups finished in
  // queue[4 s + 2]; all back to zero when the last one ends).  The first
  // nlarge entries of porder are solved by a whole workgroup of 256 threads
  // (blocks of at most 64 pfk_l entries, LDS sized by nmax_l, nblkmax_l),
  // the other nwork - nlarge by one wave each in its own wave_lds entries
  // of LDS (blocks of at most 64 pfk entries, sized by nmax, nblkmax)
  int pgrid, pthreads;
  int *queue;
  const int *porder;
  int nwork, nlarge;
  int pfk_l, nmax_l, nblkmax_l;
  int wave_lds;
};

// S_s = A_GG - W^T D^-1 W with W = L^-1 (-A_IG), all on the FP64 matrix
// cores.  W is formed over a subdomain's rows up to its last interior one
// (nsw; the rest of W is zero) and stored in chunks of 16 interface columns,
// each chunk row-major: row i of chunk c at (c nsw + i) 16.
// L_BB^-1 of the uniform 16-row blocks of the swept rows, 16x16 row-major
// (one launch per system)
struct BlockInv16Args {
  const int2 *blist;     // (subdomain, block) per block, subdomain-major
  int nblk;
  const SubMeta *meta;
  const int *nsw;        // per subdomain: rows swept
  const int *idx1, *rowbeg;
  const double2 *fac;    // this system's factors
  double2 *m16;          // [nblk][256]
};
// W by blocked forward substitution: one wavefront per (subdomain, chunk)
struct FwdArgs {
  const int2 *work;      // (subdomain, chunk) per wave; subdomain -1: padding
  int nwork;
  const SubMeta *meta;
  const int *nsw;
  const int *idx1, *rowbeg;
  const double2 *fac;
  const int8_t *bcflag;
  const int *cp_ptr, *cp_k;
  const double2 *cp_val; // this system's couplings
  const double2 *m16;
  const long long *m16_off;  // per subdomain its first block in m16
  double2 *wmat;
  const long long *w_off;
};
// one wavefront per 16x16 lower tile of -W^T D^-1 W
struct GemmArgs {
  const SubMeta *meta;
  const int4 *tiles;     // (subdomain, row block, column block, -) per wave; subdomain -1: padding
  int ntile;
  const int *nsw;
  const double2 *wmat;
  const long long *w_off;
  const double2 *dinv;   // per DOF 1/d_i
  double2 *smat;
  const long long *sm_off;
};
// A_GG added to the tiles: one workgroup per subdomain (of subs, or all)
struct AggArgs {
  const SubMeta *meta;
  const int *subs;       // the subdomains, or null: every subdomain in order
  const int *gr_ptr, *gr_col;
  const double2 *gr_val;
  const int8_t *bcflag;
  const int *kidx;
  double2 *smat;
  const long long *sm_off;
};

// one workgroup of the explicit Schur matvec: where its subdomain's tiles,
// slots and interface count are (one scalar load instead of order -> meta ->
// tile offset)
// one work item of the explicit Schur matvec: a subdomain, or a contiguous
// range of its tiles (the largest subdomains are split over several items,
// each writing its own copy of the subdomain's slots)
struct SmvWork {
  long long off;  // first tile entry of the subdomain
  int y0;         // first output slot (yslot) of this item
  int gg;         // first entry of the subdomain's interface ids (g_gid)
  int ng;         // interface DOFs
  int pq;         // index of this item's part of (p,q) (pqpart)
  int t0, t1;     // tiles [t0, t1) of the row-block-major sequence; t1 <= t0: all
};

struct SmvArgs {
  int sys0, nsys;
  SysStride ss;
  const SmvWork *work;   // per work item (batched launches), or null: order/meta/sm_off
  const SubMeta *meta;
  const int *order;
  int base;
  const int *done;
  int nbmax;             // largest number of 16-row blocks (LDS sizing)
  int waves;             // waves per workgroup: 2, 4 or 8
  const double2 *smat;
  const long long *sm_off;
  const int *g_gid;
  const double2 *P;
  const double2 *Z;      // p = Z + beta P (null: p = P)
  Scal *sc;              // status of the interface iteration (with Z)
  const double2 *part;   // (r,z) and max|r| partials of k_iter_ab (with Z)
  int G;                 // their count per system
  double2 *pqpart;       // per subdomain its part of (p,q) = sum_slots p y (explicit iteration; may be null)
  int npq;               // per system: subdomains
  double2 *yslot;
  // persistent form (pgrid > 0): pgrid workgroups per system take the work
  // items from a queue, workgroup b first item b, then the next unclaimed
  // one (queue[2 s]: items claimed past the first pgrid, queue[2 s + 1]:
  // workgroups finished; both back to zero when the last one ends)
  int pgrid;
  int *queue;
};

struct FactorArgs {
  const SubMeta *meta;
  const int *order;
  int base;
  const int *idx1;
  const int *rowbeg;
  double2 *fac;
  double2 *dinv;
  int wmax;              // LDS row buffer length
  int nmax;
  double2 *m16;          // blocked form: per subdomain and 16-row block L_BB^-1 (16x16, temporary)
  const long long *mf_off;  // per subdomain its first block in m16
};

struct AsmArgs {
  int nel;
  const int *el_vtx;     // [nel*4] global vertex index
  const double *crd;     // [nv*3]
  const double *joe;     // [nv*6]
  const double *dgv;     // [nv] delta-gap voltage per vertex, or null (current source)
  const double *eps, *sigma, *pml;
  const int *src;
  double omega;
  double2 *elem_out;     // [nel*27]: 21 lower pairs + 6 loads
  int *bad;              // set when an element is degenerate
};

// element post-processing (HDDM_post_em_E)
struct PostArgs {
  int nel;
  const int *el_vtx;     // [nel*4] global vertex index
  const int *el_dof;     // [nel*6] global DOF of each local edge
  const double *crd;     // [nv*3]
  const double *sigma;   // [nel]
  const double *nu;      // [nel] 1/mu of the element's material, or null (all 7.957747e5)
  double omega;
  const double2 *X;      // per DOF solution
  double2 *B, *E, *J;    // [nel*3] each
};

// adjoint permittivity sensitivity
struct GradArgs {
  int nel;
  const int *el_vtx;     // [nel*4]
  const int *el_dof;     // [nel*6]
  const double *crd;
  const double *pml;     // [nel*6] or null
  double omega;
  const double2 *xE;     // per DOF forward solution
  const double2 *xL;     // per DOF adjoint solution
  double *grad;          // [nel] dJ/deps
};

// scalars of the interface COCG
struct Scal {
  double2 rz;      // (r,z) of the current residual
  double2 beta;    // last beta: the next search direction is z + beta p
  double r0inv;
  double resi;
  double eps;
  int loop;
  int done;
  int conv;
  int iters;
  int max_loop;
  // the status step of an iteration ended without k_iter_c: k_iter_ab sets
  // pending; the next explicit Schur matvec takes the step from k_iter_ab's
  // partials and stages its result (st_*, st_fresh); the next k_iter_ab
  // commits it.
  // With nothing staged, st_rz, st_beta and st_done equal rz, beta and done.
  int pending;
  double2 st_rz, st_beta;
  double st_resi;
  int st_conv, st_done, st_fresh, pad;
};

// entries of VecArgs::gq_slot4: no slot, and (in .y, with SLOT_NONE in .x)
// an unknown shared by more than four subdomains
constexpr int SLOT_NONE = -2147483647 - 1;
constexpr int SLOT_MORE = -2147483647;
struct VecArgs {
  int sys0, nsys;
  SysStride ss;
  int ng;                 // N_Gamma
  const int *gq_ptr;      // gather CSR over interface unknowns
  const int *gq_slot;     // slot ids, ~slot marks the first slot of a part group
  const int4 *gq_slot4;   // the same ids, four per unknown (SLOT_NONE pads; SLOT_MORE in .y: use the CSR)
  const double2 *yslot;
  const double2 *pcslot;
  double2 *P, *Q, *R, *Z, *U, *PC;
  double2 *part;          // [3*G] partial sums
  const double2 *pqpart;  // explicit iteration: per subdomain (p,q) partials of the Schur matvec
  int npq;
  Scal *sc;
  double *hist;
  int G;                  // grid size of the vector kernels (per system)
};

// launchers (fw_kernels.hip)
hipError_t launch_element(const AsmArgs &a, hipStream_t st);
hipError_t launch_post(const PostArgs &a, hipStream_t st);
hipError_t launch_grad_eps(const GradArgs &a, hipStream_t st);
hipError_t launch_gather_sum(int n, const int *ptr, const int *src, const double2 *in,
                             double2 *out, hipStream_t st);
hipError_t launch_copy_idx(int n, const int *slot, const double2 *in, double2 *out,
                           hipStream_t st);
hipError_t launch_profile(int ndof, const int *dof_sub, const SubMeta *meta, const int *rowptr,
                          const int *col, const double2 *val, const int8_t *bcflag, int *idx1,
                          hipStream_t st);
hipError_t launch_skyfill(int nnz_tot, const int *ent_row, const int *dof_sub, const SubMeta *meta,
                          const int *col, const double2 *val, const int8_t *bcflag,
                          const int *idx1, const int *rowbeg, double2 *fac, hipStream_t st);
// identity diagonals of the Dirichlet rows among DOFs [d0, d0 + ndof)
hipError_t launch_skydiag_bc(int d0, int ndof, const int *dof_sub, const SubMeta *meta,
                             const int8_t *bcflag, const int *rowbeg, const int *idx1, double2 *fac,
                             hipStream_t st);
hipError_t launch_factor(const FactorArgs &a, int nwork, hipStream_t st);
// the blocked form on the matrix cores (fw_factor.hip)
hipError_t launch_factor16(const FactorArgs &a, int nwork, hipStream_t st);
// inverted unit lower triangles of every row block (both partitions) of
// every subdomain, from the factors
hipError_t launch_blockinv(const SubMeta *meta, int nsub, const int *idx1, const int *rowbeg, const int *blk_row,
                           const double2 *fac, double2 *minv, hipStream_t st);
hipError_t launch_solve(const SolveArgs &a, int nwork, hipStream_t st);
size_t solve_lds_bytes(int nmax, int ngmax, int nblkmax, int pfk, int waves);
// the persistent form: LDS bytes of one workgroup (the large items' layout,
// or pthreads/64 waves' own layouts, whichever is larger) and its resident
// workgroups per CU (0: does not fit)
size_t solve_persistent_lds(const SolveArgs &a);
int solve_persistent_occupancy(const SolveArgs &a);
hipError_t launch_smv(const SmvArgs &a, int nwork, hipStream_t st);
hipError_t launch_blockinv16(const BlockInv16Args &a, hipStream_t st);
hipError_t launch_schur_fwd(const FwdArgs &a, hipStream_t st);
hipError_t launch_schur_gemm(const GemmArgs &a, hipStream_t st);
hipError_t launch_schur_agg(const AggArgs &a, int nsub, hipStream_t st);
size_t smv_lds_bytes(int nbmax, int waves);
// workgroups of the persistent matvec: its resident capacity, occupancy x CUs (0: query failed)
int smv_persistent_grid(int waves, int nbmax, int ncu);
hipError_t launch_pcslot(int nslots, const int *slot_dof, const int *rowptr, const double2 *val,
                         double2 *pcslot, hipStream_t st);
hipError_t launch_first(const VecArgs &a, hipStream_t st);
// k_iter_ab and (status_step) k_iter_c: the explicit iteration after k_smv
hipError_t launch_iter_ab(const VecArgs &a, hipStream_t st, bool status_step = true);
hipError_t launch_finalize_p(const VecArgs &a, hipStream_t st);
hipError_t launch_gather_q(const VecArgs &a, hipStream_t st);
hipError_t launch_fill(double2 *p, long long n, double2 v, hipStream_t st);
hipError_t launch_scal_init(Scal *sc, int nsys, double eps, int max_loop, hipStream_t st);

}  // namespace fw

namespace fw {

// subdomain COCG with IC(0) (mymt_CS_AIJ_Solver_COCG, aij_solver_cg.h:113-355;
// mymt_CS_AIJ_PC_ICC_make/set, aij_pc_func.h:36-167)
struct IccArgs {
  const SubMeta *meta;
  const int *dptr;        // Dirichlet lower CSR, per DOF (n_dof+1, global entry ids)
  const int *dcol;        // local columns (sorted, diagonal last)
  const double2 *dval;
  double2 *L;             // IC(0) factor, same layout as dval
  double2 *dg;            // per DOF 1/L_ii
  const int *fptr;        // full symmetric rows of the Dirichlet matrix (n_dof+1)
  const int *fcol;
  const double2 *fval;
  double af;              // ICC acceleration factor
  const double2 *b;       // per DOF load
  double2 *x;             // per DOF solution (initial guess in)
  int *status;            // per subdomain: >=0 iterations, <0 reference status
  double eps, div;
  int nmax;
};
hipError_t launch_icc_make(const IccArgs &a, int nsub, hipStream_t st);
hipError_t launch_cocg(const IccArgs &a, int nsub, hipStream_t st);

}  // namespace fw
