// nrpyell_common.h -- shared host/device definitions for libnrpyell_hip.so.
//
// Grid layout is the reference's (NRPy_basic_defines.h:247-250): gridfunction-major SoA,
// i0 fastest, ghost zones of width NGHOSTS=5 stored in place.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NRPYELL_NG 5

struct GridDims {
  int N0, N1, N2;  // interior points per direction
  int G0, G1, G2;  // with ghosts
  long ntot;       // G0*G1*G2 points per gridfunction
  double invdx0, invdx1, invdx2;
};

__host__ __device__ inline long idx3(const GridDims &g, int i0, int i1, int i2) {
  return (long)i0 + (long)g.G0 * ((long)i1 + (long)g.G1 * (long)i2);
}

// Periodic index in the phi (xx2) direction: interior i2 runs over [NG, N2+NG).
__host__ __device__ inline int wrap2(const GridDims &g, int j) {
  int m = (j - NRPYELL_NG) % g.N2;
  if (m < 0) m += g.N2;
  return m + NRPYELL_NG;
}

// wrap2 for indices within one period of the interior planes, j - NG in [-N2, 2 N2): one
// conditional add and one conditional subtract instead of an integer modulo (whose generic
// sequence the compiler emits for a runtime N2).
__host__ __device__ inline int wrap2_near(const GridDims &g, int j) {
  int m = j - NRPYELL_NG;
  m += m < 0 ? g.N2 : 0;
  m -= m >= g.N2 ? g.N2 : 0;
  return m + NRPYELL_NG;
}

// floor(q / d) as a multiply-high with mul = ceil(2^32 / d) (div_mul): exact whenever
// q * d <= 2^32, which the host checks for every block index it launches.
__host__ __device__ inline unsigned div_by(unsigned q, unsigned long long mul) {
  return (unsigned)(((unsigned long long)q * mul) >> 32);
}
inline unsigned long long div_mul(unsigned d) { return ((1ull << 32) + d - 1) / d; }

// Ghost -> source map of the SinhSymTP grid, the closed form of the eigen-coordinate round
// trip of bcstruct_set_up.c:27-194:
//   xx0 < 0 : (xx0,xx1,xx2) ~ (-xx0, xx1, xx2+pi)   i0 -> 2*NG-1-i0, half turn in phi
//   xx1 < 0 : (xx0,xx1,xx2) ~ (xx0, -xx1, xx2+pi)   i1 -> 2*NG-1-i1, half turn
//   xx1 > pi: (xx0,xx1,xx2) ~ (xx0, 2pi-xx1, xx2+pi) i1 -> 2*(N1+NG)-1-i1, half turn
//   xx2 outside [-pi,pi): periodic.
// Points with i0 >= N0+NG (outside the outer radius) keep i0. A point that maps to itself is
// a pure outer point. The library checks this closed form against the bcstruct table of
// every grid it runs on (ghostmap_verify) and refuses grids where they differ.
__host__ __device__ inline void ghost_source(const GridDims &g, int i0, int i1, int i2, int &s0, int &s1,
                                             int &s2) {
  int flips = 0;
  s0 = i0;
  s1 = i1;
  if (i0 < NRPYELL_NG) { s0 = 2 * NRPYELL_NG - 1 - i0; flips++; }
  if (i1 < NRPYELL_NG) { s1 = 2 * NRPYELL_NG - 1 - i1; flips++; }
  else if (i1 >= g.N1 + NRPYELL_NG) { s1 = 2 * (g.N1 + NRPYELL_NG) - 1 - i1; flips++; }
  s2 = wrap2(g, i2 + ((flips & 1) ? g.N2 / 2 : 0));
}

__host__ __device__ inline bool is_interior(const GridDims &g, int i0, int i1, int i2) {
  return i0 >= NRPYELL_NG && i0 < g.N0 + NRPYELL_NG && i1 >= NRPYELL_NG && i1 < g.N1 + NRPYELL_NG &&
         i2 >= NRPYELL_NG && i2 < g.N2 + NRPYELL_NG;
}

// Enumerate every ghost point whose source (ghost_source) is (i0,i1,i2), an interior or pure
// outer point, calling f(dst_index). This is the exact inverse of ghost_source: candidate i0
// and i1 reflections, then every periodic phi copy shifted by the half turn their parity needs.
template <class F>
__host__ __device__ inline void for_each_image(const GridDims &g, int i0, int i1, int i2, F &&f) {
  int a_c[2], a_f[2], na = 0;
  a_c[na] = i0; a_f[na++] = 0;
  if (i0 >= NRPYELL_NG && i0 < 2 * NRPYELL_NG) { a_c[na] = 2 * NRPYELL_NG - 1 - i0; a_f[na++] = 1; }
  int b_c[3], b_f[3], nb = 0;
  b_c[nb] = i1; b_f[nb++] = 0;
  if (i1 >= NRPYELL_NG && i1 < 2 * NRPYELL_NG) { b_c[nb] = 2 * NRPYELL_NG - 1 - i1; b_f[nb++] = 1; }
  if (i1 >= g.N1 && i1 < g.N1 + NRPYELL_NG && i1 >= NRPYELL_NG) {
    b_c[nb] = 2 * (g.N1 + NRPYELL_NG) - 1 - i1; b_f[nb++] = 1;
  }
  for (int ia = 0; ia < na; ia++)
    for (int ib = 0; ib < nb; ib++) {
      const int a = a_c[ia], b = b_c[ib];
      const int shift = ((a_f[ia] + b_f[ib]) & 1) ? g.N2 / 2 : 0;
      // candidates j2 with wrap2(j2 + shift) == i2, j2 in [0, G2)
      int j2 = i2 - shift;
      while (j2 - g.N2 >= 0) j2 -= g.N2;
      while (j2 < 0) j2 += g.N2;
      for (; j2 < g.G2; j2 += g.N2) {
        const bool interior_like = (a >= NRPYELL_NG) && (b >= NRPYELL_NG && b < g.N1 + NRPYELL_NG) &&
                                   (j2 >= NRPYELL_NG && j2 < g.N2 + NRPYELL_NG);
        // (a,b,j2) with interior b and j2 and a >= NG is the point itself (interior or pure outer)
        if (interior_like) continue;
        f(idx3(g, a, b, j2));
      }
    }
}

// The images of (i0,i1,i2) that lie in the interior phi planes [NG, N2+NG) only: the i0/i1
// reflection ghosts, each in the one interior plane its half turn selects. Kernels that read
// phi neighbours through wrap2 never read the ghost planes, so the solver keeps just these
// and fills the ghost planes once, when the state is handed back (fill_ghost_planes).
template <class F>
__host__ __device__ inline void for_each_image_in_planes(const GridDims &g, int i0, int i1, int i2, F &&f) {
  const bool ra = i0 >= NRPYELL_NG && i0 < 2 * NRPYELL_NG;
  int b_c[2], nb = 0;
  if (i1 >= NRPYELL_NG && i1 < 2 * NRPYELL_NG) b_c[nb++] = 2 * NRPYELL_NG - 1 - i1;
  if (i1 >= g.N1 && i1 < g.N1 + NRPYELL_NG && i1 >= NRPYELL_NG) b_c[nb++] = 2 * (g.N1 + NRPYELL_NG) - 1 - i1;
  if (!ra && nb == 0) return;
  const int j2h = wrap2_near(g, i2 + g.N2 / 2);  // one reflection: half turn (i2 interior)
  const int a = 2 * NRPYELL_NG - 1 - i0;
  if (ra) f(idx3(g, a, i1, j2h));
  for (int ib = 0; ib < nb; ib++) {
    f(idx3(g, i0, b_c[ib], j2h));
    if (ra) f(idx3(g, a, b_c[ib], i2));  // two reflections: no turn
  }
}

// base[i] with a 32-bit element index: the address is the 64-bit (uniform) base plus a 32-bit
// byte offset, which the compiler emits as one saddr global load/store without per-lane 64-bit
// address arithmetic. Callers keep every array under 2^32 bytes. Stores are ordinary (cached):
// the next stage reads them back, and non-temporal stores, which send them past the MALL,
// measured 6% slower per launch over whole RK4 steps.
template <typename T>
__device__ __forceinline__ T ld32(const T *__restrict__ base, unsigned i) {
  return *(const T *)((const char *)base + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T *base, unsigned i, T v) {
  *(T *)((char *)base + i * (unsigned)sizeof(T)) = v;
}

// Two consecutive elements base[i], base[i+1] (i even, base 16-byte aligned) in one load.
template <typename T>
struct vec2_of;
template <>
struct vec2_of<double> {
  using type = double2;
};
template <>
struct vec2_of<float> {
  using type = float2;
};
template <typename T>
__device__ __forceinline__ void ld2x32(const T *__restrict__ base, unsigned i, T &x, T &y) {
  const typename vec2_of<T>::type v = *(const typename vec2_of<T>::type *)((const char *)base + i * (unsigned)sizeof(T));
  x = v.x;
  y = v.y;
}

// 10th-order centred FD weights (rhs_eval.c:23-27; Fornberg).
#define FD1_W1 (5.0 / 6.0)
#define FD1_W2 (5.0 / 21.0)
#define FD1_W3 (5.0 / 84.0)
#define FD1_W4 (5.0 / 504.0)
#define FD1_W5 (1.0 / 1260.0)
#define FD2_W0 (5269.0 / 1800.0)
#define FD2_W1 (5.0 / 3.0)
#define FD2_W2 (5.0 / 21.0)
#define FD2_W3 (5.0 / 126.0)
#define FD2_W4 (5.0 / 1008.0)
#define FD2_W5 (1.0 / 3150.0)

// 11-point first derivative with the stencil shifted by -s (s = 0..5; s = 0 centred); the
// Fornberg rationals of apply_bcs_outerradiation_and_inner.c:11-141. Shift +s uses
// w[j] = -c_fd1_shift[s][10-j].
static __constant__ double c_fd1_shift[6][11] = {
    {-1.0 / 1260, 5.0 / 504, -5.0 / 84, 5.0 / 21, -5.0 / 6, 0.0, 5.0 / 6, -5.0 / 21, 5.0 / 84, -5.0 / 504, 1.0 / 1260},
    {1.0 / 1260, -1.0 / 105, 3.0 / 56, -4.0 / 21, 1.0 / 2, -6.0 / 5, 11.0 / 30, 4.0 / 7, -3.0 / 28, 1.0 / 63, -1.0 / 840},
    {-1.0 / 840, 1.0 / 72, -3.0 / 40, 1.0 / 4, -7.0 / 12, 21.0 / 20, -7.0 / 4, 319.0 / 420, 3.0 / 8, -1.0 / 24, 1.0 / 360},
    {1.0 / 360, -2.0 / 63, 1.0 / 6, -8.0 / 15, 7.0 / 6, -28.0 / 15, 7.0 / 3, -8.0 / 3, 341.0 / 280, 2.0 / 9, -1.0 / 90},
    {-1.0 / 90, 1.0 / 8, -9.0 / 14, 2.0, -21.0 / 5, 63.0 / 10, -7.0, 6.0, -9.0 / 2, 4609.0 / 2520, 1.0 / 10},
    {1.0 / 10, -10.0 / 9, 45.0 / 8, -120.0 / 7, 35.0, -252.0 / 5, 105.0 / 2, -40.0, 45.0 / 2, -10.0, 7381.0 / 2520}};

// Kernels whose lanes use different rows of c_fd1_shift work from a copy of it in LDS: a
// lane-indexed read of the constant table is a vector memory load, i.e. a global-memory round
// trip wherever the stencil is evaluated. Each of the first NRPYELL_FD_TABLE threads of a block
// loads one entry with its other loads and stores it before the block's first barrier.
#define NRPYELL_FD_TABLE 66
__device__ __forceinline__ double load_fd_table(int tid) {
  return tid < NRPYELL_FD_TABLE ? (&c_fd1_shift[0][0])[tid] : 0.0;
}

// 7-point one-sided first-derivative weights (6th order; taps 0..6 from the point outward):
// -49/20, 6, -15/2, 20/3, -15/4, 6/5, -1/6.
static __constant__ double c_fd1_onesided7[7] = {-49.0 / 20.0, 6.0, -15.0 / 2.0, 20.0 / 3.0,
                                                 -15.0 / 4.0,  6.0 / 5.0, -1.0 / 6.0};
