// fw_device.h - device helpers shared by the libfwhip kernels (gfx950).
//
// Complex FP64 values are double2 (re, im).  Products follow the reference's
// ComplexMulti (lib/std/complex_adv.c:103-114): re = ac - bd, im = bc + ad.
#pragma once
#include <hip/hip_runtime.h>

namespace fw {

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 acc, double2 a, double2 b) {
  return make_double2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.y * b.x + a.x * b.y));
}
// acc - a*b
__device__ __forceinline__ double2 cfms(double2 acc, double2 a, double2 b) {
  return make_double2(acc.x - (a.x * b.x - a.y * b.y), acc.y - (a.y * b.x + a.x * b.y));
}
// z + b*p with explicit fused multiply-adds, so that every kernel that forms
// the next search direction obtains the same bits
__device__ __forceinline__ double2 zbp(double2 z, double2 b, double2 p) {
  return make_double2(__fma_rn(b.x, p.x, __fma_rn(-b.y, p.y, z.x)), __fma_rn(b.x, p.y, __fma_rn(b.y, p.x, z.y)));
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabsv(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ bool ciszero(double2 a) { return a.x == 0.0 && a.y == 0.0; }

// naive reciprocal of the reference's ComplexReciprocal
// (lib/std/complex_adv.c:39-51): scalar interface updates use it.
__device__ __forceinline__ double2 crecip_naive(double2 a) {
  double d = 1.0 / (a.x * a.x + a.y * a.y);
  return make_double2(a.x * d, -(a.y * d));
}
__device__ __forceinline__ double2 cdiv_naive(double2 a, double2 b) { return cmul(a, crecip_naive(b)); }

// 1/d by R. L. Smith's algorithm (Comm. ACM 5(8), 1962): scaled by the
// larger component, so it neither overflows nor underflows for pivots of
// any magnitude.  Used for the skyline pivots, where the reference relies on
// the C99 complex division of its host compiler.
__device__ __forceinline__ double2 crecip_smith(double2 d) {
  if (fabs(d.x) >= fabs(d.y)) {
    double r = d.y / d.x;
    double den = d.x + d.y * r;
    return make_double2(1.0 / den, -r / den);
  } else {
    double r = d.x / d.y;
    double den = d.y + d.x * r;
    return make_double2(r / den, -1.0 / den);
  }
}

// ---- wavefront (64-lane) reductions --------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double2 wave_csum(double2 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
constexpr double NEARLY_ZERO = 1.0e-60;  // lib/std/defs.h:105

// The status step of the interface COCG after an iteration: beta from the new
// and old (r,z) (hddm_solver_s.h:340-350, both near zero: 0), the max-norm
// residual ratio and stop test (:311-318), and the status machine's end of
// the loop on convergence or at max_loop (hddm_solver.h:125-143).
struct StatusStep {
  double2 beta;
  double resi;
  int conv, done;
};
__device__ __forceinline__ StatusStep status_step(double2 rzs, double mxs, double2 rzo, double r0inv, double eps,
                                                  int loop, int max_loop) {
  StatusStep r;
  r.resi = mxs * r0inv;
  r.beta = cz();
  if (!(cabsv(rzs) < NEARLY_ZERO && cabsv(rzo) < NEARLY_ZERO)) r.beta = cdiv_naive(rzs, rzo);
  r.conv = r.resi < eps ? 1 : 0;
  r.done = (r.conv || loop == max_loop) ? 1 : 0;
  return r;
}

// Fixed-order reductions of G per-block partials, as a 256-thread block
// computes them (thread t sums entries t, t+256, ... in order, each 64-lane
// wave sums its threads, then the four wave sums are added in order), with
// any number NT of threads: wave w takes the virtual waves w, w+NT/64, ... < 4.
// Every kernel that needs the same scalar thus obtains the same bits.
// sh: 4 entries of LDS; contains barriers (call uniformly).
__device__ __forceinline__ double2 parts_csum256(const double2 *__restrict__ p, int G, double2 *sh, int NT) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int vw = wave; vw < 4; vw += NT >> 6) {
    double2 t = cz();
    for (int k = vw * 64 + lane; k < G; k += 1024) {  // four loads in flight, added in order
      double2 v[4];
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] = k + 256 * j < G ? p[k + 256 * j] : cz();
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k + 256 * j < G) t = cadd(t, v[j]);
    }
    t = wave_csum(t);
    if (lane == 0) sh[vw] = t;
  }
  __syncthreads();
  double2 t = cz();
#pragma unroll
  for (int k = 0; k < 4; k++) t = cadd(t, sh[k]);
  return t;
}
__device__ __forceinline__ double parts_max256(const double2 *__restrict__ p, int G, double *sh, int NT) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int vw = wave; vw < 4; vw += NT >> 6) {
    double t = 0.0;
    for (int k = vw * 64 + lane; k < G; k += 1024) {
      double v[4];
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] = k + 256 * j < G ? p[k + 256 * j].x : 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k + 256 * j < G) t = fmax(t, v[j]);
    }
    t = wave_max(t);
    if (lane == 0) sh[vw] = t;
  }
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int k = 1; k < 4; k++) t = fmax(t, sh[k]);
  return t;
}

// sum over aligned groups of G lanes (G power of two <= 64)
template <int G>
__device__ __forceinline__ double2 group_csum(double2 v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, G);
    v.y += __shfl_xor(v.y, o, G);
  }
  return v;
}

// broadcast lane `lane` (wave-uniform) of a double through the scalar unit
__device__ __forceinline__ double readlane_d(double x, int lane) {
  long long b = __double_as_longlong(x);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double2 readlane_c(double2 x, int lane) {
  return make_double2(readlane_d(x.x, lane), readlane_d(x.y, lane));
}

// a double moved between lanes by a DPP pattern (two 32-bit moves; every
// lane active, patterns inside a 16-lane row)
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// the sum over each aligned group of 2^lsh lanes (lsh = 2..6) in every lane
// of the group: quad_perm [1,0,3,2] and [2,3,0,1] (the quad sums),
// row_half_mirror (adds the other quad of the 8-lane half-row),
// row_mirror (the other half of the 16-lane row), then exchanges across
// 16 and 32 lanes.  Each step adds the two partial sums of a pair of
// disjoint lane sets, so every lane of a group ends with the same bits, the
// bits of the xor butterfly over the same group.
__device__ __forceinline__ double group_sum(double v, int lsh) {
  v += dpp_mov_d<0xB1>(v);
  v += dpp_mov_d<0x4E>(v);
  if (lsh >= 3) v += dpp_mov_d<0x141>(v);
  if (lsh >= 4) v += dpp_mov_d<0x140>(v);
  if (lsh >= 5) v += __shfl_xor(v, 16, 64);
  if (lsh >= 6) v += __shfl_xor(v, 32, 64);
  return v;
}
__device__ __forceinline__ double2 group_csum(double2 v, int lsh) {
  return make_double2(group_sum(v.x, lsh), group_sum(v.y, lsh));
}

}  // namespace fw
