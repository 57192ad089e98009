// nrpyell_stencil.h -- device pieces shared by the fused substep kernels (one RK4 substep per
// launch, nrpyell_substep.hip; two substeps per launch, nrpyell_substep2.hip): the RK4 stage
// updates, the 10th-order stencil sums, the RHS at one point, block sums and the XCD-aware
// block order.
#pragma once
#include <hip/hip_runtime.h>

#include "nrpyell_common.h"
#include "nrpyell_internal.h"

// Stage updates, MoL_step_forward_in_time.c:44-45 (k1), :88-89 (k2), :129-130 (k3), :170 (k4).
template <typename T, int STAGE>
__device__ __forceinline__ void stage_update(T k, T dt, T yn, T &rt, T &out) {
  if (STAGE == 1) {
    rt = (T)(1.0 / 6.0) * (dt * k);
    out = fma((T)0.5, dt * k, yn);
  } else if (STAGE == 2) {
    rt = fma((T)(1.0 / 3.0), dt * k, rt);
    out = fma((T)0.5, dt * k, yn);
  } else if (STAGE == 3) {
    rt = fma((T)(1.0 / 3.0), dt * k, rt);
    out = fma(dt, k, yn);
  } else {
    out = yn + fma((T)(1.0 / 6.0), dt * k, rt);
  }
}

// Raw 10th-order second-difference and first-difference sums of one direction
// (weights rhs_eval.c:23-27; 1/dx factors are in the coefficient table).
template <typename T>
__device__ __forceinline__ void fd_sums(const T m[6], const T p[6], T uu, T &s2, T &s1) {
  s2 = -(T)FD2_W0 * uu;
  s2 = fma((T)FD2_W1, m[1] + p[1], s2);
  s2 = fma((T)-FD2_W2, m[2] + p[2], s2);
  s2 = fma((T)FD2_W3, m[3] + p[3], s2);
  s2 = fma((T)-FD2_W4, m[4] + p[4], s2);
  s2 = fma((T)FD2_W5, m[5] + p[5], s2);
  s1 = (T)FD1_W1 * (p[1] - m[1]);
  s1 = fma((T)-FD1_W2, p[2] - m[2], s1);
  s1 = fma((T)FD1_W3, p[3] - m[3], s1);
  s1 = fma((T)-FD1_W4, p[4] - m[4], s1);
  s1 = fma((T)FD1_W5, p[5] - m[5], s1);
}

// Fixed-order sum of the NW per-wave partials of a block (reads s[0..NW-1] only).
template <int NW>
__device__ __forceinline__ double wave_sum(const double *s) {
  static_assert(NW == 2 || NW == 4 || NW == 8, "blocks of 128, 256 or 512 threads");
  if constexpr (NW == 2) {
    return s[0] + s[1];
  } else {
    double acc = (s[0] + s[1]) + (s[2] + s[3]);
    if constexpr (NW == 8) acc += (s[4] + s[5]) + (s[6] + s[7]);
    return acc;
  }
}

// RHS at one point (rhs_eval.c:86-110): ku = v - eta u and kv = c^2 (lap u + AA^2/(8 (psi+u)^7))
// from the stencil arms, with c^2 folded into K and AAc. The interior blocks and the BC blocks
// both evaluate it at the last interior point, from the same inputs: bit-identical values.
template <typename T>
__device__ __forceinline__ void point_rhs(T uu, T vv, T psi, T AAc, const T K[5], const T xm[6], const T xp[6],
                                          const T ym[6], const T yp[6], const T zm[6], const T zp[6], T eta, T &ku,
                                          T &kv) {
  T s00, s0, s11, s1d, s22, s2d;
  fd_sums<T>(xm, xp, uu, s00, s0);
  fd_sums<T>(ym, yp, uu, s11, s1d);
  fd_sums<T>(zm, zp, uu, s22, s2d);
  const T X = psi + uu, X2 = X * X, X4 = X2 * X2, X7 = X4 * X2 * X;
  kv = AAc / X7;
  kv = fma(K[0], s00, kv);
  kv = fma(K[1], s11, kv);
  kv = fma(K[2], s22, kv);
  kv = fma(K[3], s1d, kv);
  kv = fma(K[4], s0, kv);
  ku = fma(-eta, uu, vv);
}

// XCD-aware block order. Workgroups are dispatched to the 8 XCDs round robin (workgroup L to
// XCD L % 8), and each XCD has its own L2. Logical block q = (L % 8) * (n / 8) + L / 8 gives
// every XCD one contiguous run of logical blocks; with the i2 plane group fastest in q, the
// blocks that read a tile column's i2 arm, and the neighbouring tiles that read its i1 halos,
// share one L2. Blocks past the last multiple of 8 keep their index.
__device__ __forceinline__ unsigned xcd_logical(unsigned L, unsigned n) {
  const unsigned per = n / 8;
  return L < 8 * per ? (L % 8) * per + L / 8 : L;
}

