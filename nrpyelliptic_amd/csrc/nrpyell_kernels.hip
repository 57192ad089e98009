// nrpyell_kernels.hip -- gfx950 kernels of the hyperbolic-relaxation hot path other than the fused
// RK4 substeps (nrpyell_substep.hip, nrpyell_substep2.hip). One thread per grid point or list
// entry, i0 across lanes so that loads coalesce:
//   rhs_exact_kernel, residual_exact_kernel   rhs_eval.c:7-117, residual_all_points.c:7-112 with
//                           the reference FMA tree (drop-in entry points; 64 x 4 tiles, x0/x1
//                           stencil arms staged in LDS)
//   coef_fold_kernel, fold_source_kernel      the solver's tables from rfm and auxevol
//   residual_l2_partial_kernel, l2_partial_kernel, l2_finalize_kernel   residual + masked,
//                           volume-weighted sum of squares, partials reduced in a fixed order
//                           (L2_norm_of_gf.c:6-43)
//   radiation_rows_kernel, radiation_list_kernel, sommerfeld_list_kernel, inner_gather_kernel
//                           drop-in apply_bcs_* entry points and the thorn's Sommerfeld BC
//   rfm_kernel, auxevol_kernel, wavespeed_kernel   set-up producers
//   fill_ghost_planes_kernel, settle_fill_kernel, convert, fill, shift_v   state handling
//
// Built with -ffp-contract=off: every fused multiply-add below is an explicit fma().
#include <hip/hip_runtime.h>
#include <math.h>

#include "nrpyell_common.h"
#include "nrpyell_internal.h"

#define NG NRPYELL_NG

// ---------------------------------------------------------------------------------------
// Reference-exact interior operator (rhs_eval.c:86-110; SIMD_intrinsics.h FMA semantics).
// ---------------------------------------------------------------------------------------
struct RfmPtrs {
  const double *f0, *f0_D0, *f0_DD00, *f1, *f1_D1, *f2, *f2_D0, *f3, *f3_D0;
};

__device__ inline double fd1_exact(const double m[6], const double p[6]) {
  return fma(5.0 / 504.0, m[4] - p[4],
             fma(5.0 / 6.0, p[1] - m[1],
                 fma(5.0 / 84.0, p[3] - m[3], fma(1.0 / 1260.0, p[5] - m[5], (5.0 / 21.0) * (m[2] - p[2])))));
}
__device__ inline double fd2_exact(const double m[6], const double p[6], double c0) {
  return fma(5.0 / 126.0, m[3] + p[3],
             fma(5.0 / 3.0, m[1] + p[1],
                 fma(1.0 / 3150.0, m[5] + p[5], -fma(5.0 / 1008.0, m[4] + p[4], fma(5.0 / 21.0, m[2] + p[2], c0)))));
}

// Flat-space Laplacian in SinhSymTP plus AA^2/(8 (psi+u)^7), reference expression tree, from the
// stencil arms m*/p* (u at offsets -k / +k along x0, x1, x2) and the centre value uu.
__device__ inline double lap_source_exact(const GridDims &g, const RfmPtrs &r, const double m0[6], const double p0[6],
                                          const double m1[6], const double p1[6], const double m2[6],
                                          const double p2[6], double uu, int i0, int i1, double psi, double AA) {
  const double c0 = (5269.0 / 1800.0) * uu;
  const double dD0 = g.invdx0 * fd1_exact(m0, p0);
  const double dD1 = g.invdx1 * fd1_exact(m1, p1);
  const double dDD00 = (g.invdx0 * g.invdx0) * fd2_exact(m0, p0, c0);
  const double dDD11 = (g.invdx1 * g.invdx1) * fd2_exact(m1, p1, c0);
  const double dDD22 = (g.invdx2 * g.invdx2) * fd2_exact(m2, p2, c0);
  const long k2 = i0 + (long)g.G0 * i1;
  const double f0 = r.f0[i0], f0p = r.f0_D0[i0], f0pp = r.f0_DD00[i0];
  const double f1 = r.f1[i1], f1p = r.f1_D1[i1];
  const double f2 = r.f2[k2], f2_0 = r.f2_D0[k2];
  const double f3 = r.f3[i0], f3p = r.f3_D0[i0];
  const double a0 = 1.0 / (f2 * f2);
  const double a2 = (f3 * f3) / (f0p * f0p);
  const double a3 = 1.0 / ((f2 * f2) * f2);
  const double X = psi + uu;
  const double X7 = (((((X * X) * X) * X) * X) * X) * X;
  const double num = fma(f3, fma(f0p, f2_0, f0pp * f2), -(f0p * (f2 * f3p)));
  const double tA = num / ((f0p * f0p) * f0p);
  const double tB = ((a0 * -1.0) * ((f3 * f3) / f0p)) / f0;
  const double inner = fma(a2, a3 * f2_0, -tB);
  const double B0 = fma(a3 * f3, tA, -inner);
  double e = fma(a0, dDD11, -(dD0 * B0));
  e = fma(a0 * f1p, dD1 / f1, e);
  e = fma(1.0 / 8.0, AA / X7, e);
  e = fma(a0, a2 * dDD00, e);
  e = fma(dDD22, (1.0 / (f1 * f1)) / (f0 * f0), e);
  return e;
}

// The drop-in kernels work on 64 x 4 tiles of one plane: the tile's u with its 5-point halos in
// x0 and x1 is staged in LDS (the x0 and x1 arms of every point are read from there), the x2
// arms come from global memory. Only where the values come from differs from a per-point
// gather: each point's arithmetic is the reference's expression tree.
#define EX_TX 64
#define EX_TY 4
#define EX_BX (EX_TX + 2 * NG)
#define EX_BY (EX_TY + 2 * NG)

struct ExactTile {
  int i0, i1, i2;
  long p;
  bool valid;
};

// Stage the tile's box (rows i1s-5 .. i1s+8, columns i0s-5 .. i0s+68, clamped to the grid: a
// clamped value is never read by a point of the grid) and wait for the whole block.
__device__ __forceinline__ ExactTile exact_tile(const GridDims &g, const double *u, double *box) {
  const int i0s = NG + blockIdx.x * EX_TX, i1s = NG + blockIdx.y * EX_TY, i2 = NG + blockIdx.z;
  const long base = (long)i2 * g.G0 * g.G1;
  for (int k = threadIdx.y * EX_TX + threadIdx.x; k < EX_BX * EX_BY; k += EX_TX * EX_TY) {
    const int bx = k % EX_BX, by = k / EX_BX;
    const int j0 = min(i0s - NG + bx, g.G0 - 1), j1 = min(i1s - NG + by, g.G1 - 1);
    box[k] = u[base + (long)j1 * g.G0 + j0];
  }
  __syncthreads();
  ExactTile t;
  t.i0 = i0s + (int)threadIdx.x;
  t.i1 = i1s + (int)threadIdx.y;
  t.i2 = i2;
  t.p = idx3(g, t.i0, t.i1, i2);
  t.valid = t.i0 < g.N0 + NG && t.i1 < g.N1 + NG;
  return t;
}

__device__ __forceinline__ double exact_point(const GridDims &g, const RfmPtrs &r, const double *u, const double *box,
                                              const ExactTile &t, double psi, double AA) {
  double m0[6], p0[6], m1[6], p1[6], m2[6], p2[6];
  const int c = ((int)threadIdx.y + NG) * EX_BX + (int)threadIdx.x + NG;
  const long s2 = (long)g.G0 * g.G1;
#pragma unroll
  for (int k = 1; k <= 5; k++) {
    m0[k] = box[c - k]; p0[k] = box[c + k];
    m1[k] = box[c - k * EX_BX]; p1[k] = box[c + k * EX_BX];
    m2[k] = u[t.p - k * s2]; p2[k] = u[t.p + k * s2];
  }
  return lap_source_exact(g, r, m0, p0, m1, p1, m2, p2, box[c], t.i0, t.i1, psi, AA);
}

__global__ void __launch_bounds__(256) rhs_exact_kernel(GridDims g, RfmPtrs r, const double *aux, const double *in,
                                                        double *out, double eta) {
  __shared__ double box[EX_BX * EX_BY];
  const ExactTile t = exact_tile(g, in, box);
  if (!t.valid) return;
  const long p = t.p, nt = g.ntot;
  const double *u = in, *v = in + nt;
  const double e = exact_point(g, r, u, box, t, aux[PSI_BACKGROUNDGF * nt + p], aux[ADD_TIMES_AUUGF * nt + p]);
  const double c = aux[WAVESPEEDGF * nt + p];
  out[p] = fma(-eta, u[p], v[p]);
  out[nt + p] = (c * c) * e;
}

__global__ void __launch_bounds__(256) residual_exact_kernel(GridDims g, RfmPtrs r, const double *aux,
                                                             const double *in, double *out) {
  __shared__ double box[EX_BX * EX_BY];
  const ExactTile t = exact_tile(g, in, box);
  if (!t.valid) return;
  const long p = t.p, nt = g.ntot;
  out[p] = exact_point(g, r, in, box, t, aux[PSI_BACKGROUNDGF * nt + p], aux[ADD_TIMES_AUUGF * nt + p]);
}

void launch_rhs_exact(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *in, double *out,
                      double eta, hipStream_t st) {
  RfmPtrs r{rfm->f0_of_xx0, rfm->f0_of_xx0__D0, rfm->f0_of_xx0__DD00, rfm->f1_of_xx1, rfm->f1_of_xx1__D1,
            rfm->f2_of_xx0_xx1, rfm->f2_of_xx0_xx1__D0, rfm->f3_of_xx0, rfm->f3_of_xx0__D0};
  dim3 blk(EX_TX, EX_TY), grd((g.N0 + EX_TX - 1) / EX_TX, (g.N1 + EX_TY - 1) / EX_TY, g.N2);
  if (out)
    hipLaunchKernelGGL(rhs_exact_kernel, grd, blk, 0, st, g, r, aux, in, out, eta);
}

void launch_residual_exact(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *in,
                           double *out, hipStream_t st) {
  RfmPtrs r{rfm->f0_of_xx0, rfm->f0_of_xx0__D0, rfm->f0_of_xx0__DD00, rfm->f1_of_xx1, rfm->f1_of_xx1__D1,
            rfm->f2_of_xx0_xx1, rfm->f2_of_xx0_xx1__D0, rfm->f3_of_xx0, rfm->f3_of_xx0__D0};
  dim3 blk(EX_TX, EX_TY), grd((g.N0 + EX_TX - 1) / EX_TX, (g.N1 + EX_TY - 1) / EX_TY, g.N2);
  hipLaunchKernelGGL(residual_exact_kernel, grd, blk, 0, st, g, r, aux, in, out);
}

// ---------------------------------------------------------------------------------------
// Precombined metric coefficients for the fused solver: the (i0,i1)-dependent factors of the
// reference expression (rhs_eval.c:98) with the FD 1/dx factors folded in, so the per-point
// work is 5 FMAs and one division:
//   L u = K00 s00 + K11 s11 + K22 s22 + K1 s1 + K0 s0
//   s00 = sum_k w2_k (u[i+k]+u[i-k]) - w2_0 u  (raw second difference along x0; s11, s22 alike)
//   s0  = sum_k w1_k (u[i+k]-u[i-k])           (raw first difference along x0; s1 alike)
//   K00 = f3^2/(f0'^2 f2^2) dx0^-2, K11 = dx1^-2/f2^2, K22 = dx2^-2/(f0^2 f1^2),
//   K1  = f1'/(f1 f2^2) dx1^-1,     K0 = -B dx0^-1,
//   B   = f3 (f3 (f0' d0f2 + f0'' f2) - f0' f2 f3')/(f0'^3 f2^3) - d0f2 f3^2/(f0'^2 f2^3)
//         - f3^2/(f0 f0' f2^2)
// ---------------------------------------------------------------------------------------
// coef_fold_kernel writes K00 .. K0 (coef, read by the residual check) and the same five times
// c^2 plus the L2 weight / c^4 (kc64, kc32: the fused kernels' per-(i0,i1) record, c the wave
// speed on the plane i2 = NG), and zeroes the flags it is given (thread 0 of block (0, 0)):
// the wave-speed check's, which fold_source_kernel then raises, and the caller's.
__global__ void coef_fold_kernel(GridDims g, RfmPtrs r, const double *aux, const double *w2d, double *coef,
                                 double *kc64, float *kc32, int *flag0, int *flag1) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = blockIdx.y;
  if (i0 == 0 && i1 == 0) {
    *flag0 = 0;
    if (flag1) *flag1 = 0;
  }
  if (i0 >= g.G0) return;
  const long k2 = i0 + (long)g.G0 * i1, n2 = (long)g.G0 * g.G1;
  const double f0 = r.f0[i0], f0p = r.f0_D0[i0], f0pp = r.f0_DD00[i0];
  const double f1 = r.f1[i1], f1p = r.f1_D1[i1];
  const double f2 = r.f2[k2], f2_0 = r.f2_D0[k2];
  const double f3 = r.f3[i0], f3p = r.f3_D0[i0];
  const double a0 = 1.0 / (f2 * f2), a2 = (f3 * f3) / (f0p * f0p), a3 = a0 / f2;
  const double B = f3 * a3 * (f3 * (f0p * f2_0 + f0pp * f2) - f0p * f2 * f3p) / (f0p * f0p * f0p) -
                   a2 * a3 * f2_0 - a0 * (f3 * f3) / (f0p * f0);
  const double K[5] = {a0 * a2 * (g.invdx0 * g.invdx0), a0 * (g.invdx1 * g.invdx1),
                       (1.0 / (f1 * f1)) / (f0 * f0) * (g.invdx2 * g.invdx2), a0 * f1p / f1 * g.invdx1,
                       -B * g.invdx0};
  const double c = aux[WAVESPEEDGF * g.ntot + idx3(g, i0, i1, NG)], c2 = c * c;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    coef[k * n2 + k2] = K[k];
    const double v = c2 * K[k];
    kc64[6 * k2 + k] = v;
    kc32[6 * k2 + k] = (float)v;
  }
  const double w = c2 > 0.0 ? w2d[k2] / (c2 * c2) : 0.0;
  kc64[6 * k2 + 5] = w;
  kc32[6 * k2 + 5] = (float)w;
}

// (psi, c^2 AA^2 / 8) interleaved per point (one 16-byte load in the fused kernels); raises *bad
// where the wave speed at an interior point differs from its value on the plane i2 = NG.
__global__ void fold_source_kernel(GridDims g, const double *aux, double *src, int *bad) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x, nt = g.ntot;
  if (p >= nt) return;
  const int plane = g.G0 * g.G1;
  const int i2 = (int)(p / plane), rem = (int)(p - (long)i2 * plane), i1 = rem / g.G0, i0 = rem - i1 * g.G0;
  const double c = aux[WAVESPEEDGF * nt + p];
  src[2 * p] = aux[PSI_BACKGROUNDGF * nt + p];
  src[2 * p + 1] = 0.125 * (c * c) * aux[ADD_TIMES_AUUGF * nt + p];
  const bool inside = i0 >= NG && i0 < g.N0 + NG && i1 >= NG && i1 < g.N1 + NG && i2 >= NG && i2 < g.N2 + NG;
  if (inside && c != aux[WAVESPEEDGF * nt + (p - (long)(i2 - NG) * plane)]) *bad = 1;
}

void launch_tables(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *w2d, double *coef64,
                   double *kc64, float *kc32, double *src, int *bad, int *also_clear, hipStream_t st) {
  RfmPtrs r{rfm->f0_of_xx0, rfm->f0_of_xx0__D0, rfm->f0_of_xx0__DD00, rfm->f1_of_xx1, rfm->f1_of_xx1__D1,
            rfm->f2_of_xx0_xx1, rfm->f2_of_xx0_xx1__D0, rfm->f3_of_xx0, rfm->f3_of_xx0__D0};
  hipLaunchKernelGGL(coef_fold_kernel, dim3((g.G0 + 63) / 64, g.G1), dim3(64), 0, st, g, r, aux, w2d, coef64, kc64,
                     kc32, bad, also_clear);
  hipLaunchKernelGGL(fold_source_kernel, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, g, aux, src, bad);
}

// ---------------------------------------------------------------------------------------
// Fused RK4 substep.
// ---------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------
// Fused residual + L2 (residual_all_points.c + L2_norm_of_gf.c): per point
// w(i0,i1) * R^2 where w = dV inside the integration radius and 0 outside; each block writes
// one partial sum, then one block adds the partials in index order (deterministic), and
// finishes log10(1e-16 + sqrt(S/V)). With a `state` pointer it also advances the solver's
// device-side convergence state (step counter, done flag, history).
// ---------------------------------------------------------------------------------------
template <typename TIN>
__global__ void __launch_bounds__(256) residual_l2_partial_kernel(GridDims g, const TIN *u, const double *aux,
                                                                  const double *coef, const double *w2d,
                                                                  double *partial, const int *done,
                                                                  double *res_out) {
  if (done && *done) return;
  const int i0 = NG + blockIdx.x * 64 + (threadIdx.x % 64);
  const int i1 = NG + blockIdx.y * 4 + threadIdx.x / 64;
  const int i2 = NG + blockIdx.z;
  double acc = 0.0;
  if (i0 < g.N0 + NG && i1 < g.N1 + NG) {
    const long p = idx3(g, i0, i1, i2), nt = g.ntot, n2 = (long)g.G0 * g.G1, k2 = i0 + (long)g.G0 * i1;
    const double w = w2d[k2];
    double K[5];
#pragma unroll
    for (int k = 0; k < 5; k++) K[k] = coef[k * n2 + k2];
    double m0[6], p0[6], m1[6], p1[6], m2[6], p2[6];
    const long s1 = g.G0, s2 = n2;
#pragma unroll
    for (int k = 1; k <= 5; k++) {
      m0[k] = (double)u[p - k]; p0[k] = (double)u[p + k];
      m1[k] = (double)u[p - k * s1]; p1[k] = (double)u[p + k * s1];
      // phi neighbours through the periodic map (the solver leaves the ghost planes stale)
      m2[k] = (double)u[(long)wrap2(g, i2 - k) * s2 + k2]; p2[k] = (double)u[(long)wrap2(g, i2 + k) * s2 + k2];
    }
    const double uu = (double)u[p];
    double s00 = -FD2_W0 * uu, s11 = s00, s22 = s00;
    s00 = fma(FD2_W1, m0[1] + p0[1], s00); s11 = fma(FD2_W1, m1[1] + p1[1], s11); s22 = fma(FD2_W1, m2[1] + p2[1], s22);
    s00 = fma(-FD2_W2, m0[2] + p0[2], s00); s11 = fma(-FD2_W2, m1[2] + p1[2], s11); s22 = fma(-FD2_W2, m2[2] + p2[2], s22);
    s00 = fma(FD2_W3, m0[3] + p0[3], s00); s11 = fma(FD2_W3, m1[3] + p1[3], s11); s22 = fma(FD2_W3, m2[3] + p2[3], s22);
    s00 = fma(-FD2_W4, m0[4] + p0[4], s00); s11 = fma(-FD2_W4, m1[4] + p1[4], s11); s22 = fma(-FD2_W4, m2[4] + p2[4], s22);
    s00 = fma(FD2_W5, m0[5] + p0[5], s00); s11 = fma(FD2_W5, m1[5] + p1[5], s11); s22 = fma(FD2_W5, m2[5] + p2[5], s22);
    double s0 = FD1_W1 * (p0[1] - m0[1]), s1d = FD1_W1 * (p1[1] - m1[1]);
    s0 = fma(-FD1_W2, p0[2] - m0[2], s0); s1d = fma(-FD1_W2, p1[2] - m1[2], s1d);
    s0 = fma(FD1_W3, p0[3] - m0[3], s0); s1d = fma(FD1_W3, p1[3] - m1[3], s1d);
    s0 = fma(-FD1_W4, p0[4] - m0[4], s0); s1d = fma(-FD1_W4, p1[4] - m1[4], s1d);
    s0 = fma(FD1_W5, p0[5] - m0[5], s0); s1d = fma(FD1_W5, p1[5] - m1[5], s1d);
    const double psi = aux[PSI_BACKGROUNDGF * nt + p], AA = aux[ADD_TIMES_AUUGF * nt + p];
    const double X = psi + uu, X2 = X * X, X4 = X2 * X2, X7 = X4 * X2 * X;
    double e = 0.125 * (AA / X7);
    e = fma(K[0], s00, e);
    e = fma(K[1], s11, e);
    e = fma(K[2], s22, e);
    e = fma(K[3], s1d, e);
    e = fma(K[4], s0, e);
    if (res_out) res_out[p] = e;
    acc = w * e * e;
  }
  // block reduction: wave64 shuffles then LDS across the 4 waves
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    partial[b] = ((sw[0] + sw[1]) + (sw[2] + sw[3]));
  }
}

// L2 of an arbitrary gridfunction (drop-in L2_norm_of_gf): w * gf^2 partials.
__global__ void __launch_bounds__(256) l2_partial_kernel(GridDims g, const double *gf, const double *w2d,
                                                         double *partial) {
  const int i0 = NG + blockIdx.x * 64 + (threadIdx.x % 64);
  const int i1 = NG + blockIdx.y * 4 + threadIdx.x / 64;
  const int i2 = NG + blockIdx.z;
  double acc = 0.0;
  if (i0 < g.N0 + NG && i1 < g.N1 + NG) {
    const double v = gf[idx3(g, i0, i1, i2)];
    acc = w2d[i0 + (long)g.G0 * i1] * v * v;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    partial[b] = ((sw[0] + sw[1]) + (sw[2] + sw[3]));
  }
}

// Fixed-order sum of the partials by one block; writes log10(1e-16+sqrt(S/V)) and, when `st`
// is given, updates the solver convergence state.
__global__ void __launch_bounds__(256) l2_finalize_kernel(const double *partial, long npart, double volume,
                                                          SolverDeviceState *st, int steps_per_check,
                                                          double target, double *log10_out) {
  if (st && st->done) return;
  double acc = 0.0;
  for (long i = threadIdx.x; i < npart; i += 256) acc += partial[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double S = (sw[0] + sw[1]) + (sw[2] + sw[3]);
    const double l = log10(1e-16 + sqrt(S / volume));
    if (log10_out) *log10_out = l;
    if (st) {
      st->steps += steps_per_check;
      record_check(st, l, target);
    }
  }
}

// L2 volume weights on the device (the allocation-free L2 entry point): per (i0,i1) the dV of
// host_l2_weights from the coordinate values xx0[i0], xx1[i1], zero outside the interior and
// outside the integration radius. Same expression order as the host; exp/sin/cos are the
// device's, so a weight can differ from the host's in the last bit.
__global__ void __launch_bounds__(256) l2_weights_kernel(GridDims g, L2WeightParams q, const double *xx0,
                                                         const double *xx1, double *w2d) {
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)g.G0 * g.G1) return;
  const int i0 = (int)(k % g.G0), i1 = (int)(k / g.G0);
  double dV = 0.0;
  if (i0 >= NG && i0 < NG + g.N0 && i1 >= NG && i1 < NG + g.N1) {
    const double w = q.SINHWAA, A = q.AMAX, b = q.bScale;
    const double D = exp(1.0 / w) - exp(-1.0 / w);
    const double ep = exp(xx0[i0] / w), em = exp(-xx0[i0] / w);
    const double f0 = A * (ep - em) / D, f0p = A * (ep + em) / (w * D);
    const double s = sin(xx1[i1]), c = cos(xx1[i1]);
    const double f3sq = f0 * f0 + b * b, f2sq = f0 * f0 + b * b * s * s;
    const double r = sqrt(s * s * f0 * f0 + f3sq * c * c);
    if (r < q.radius) dV = f0 * f0p * f2sq * fabs(s) / sqrt(f3sq) * q.dxx0 * q.dxx1 * q.dxx2;
  }
  w2d[k] = dV;
}

// Fixed-order sum of the partials and of the weight table by one block:
// log10(1e-16 + sqrt(S / V)) with V = N2 * sum(w2d) (every i2 plane carries the same weights).
__global__ void __launch_bounds__(256) l2_finalize_weights_kernel(const double *partial, long npart,
                                                                  const double *w2d, long nw, int n2,
                                                                  double *log10_out) {
  double acc = 0.0, vol = 0.0;
  for (long i = threadIdx.x; i < npart; i += 256) acc += partial[i];
  for (long i = threadIdx.x; i < nw; i += 256) vol += w2d[i];
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_down(acc, o, 64);
    vol += __shfl_down(vol, o, 64);
  }
  __shared__ double sw[4], sv[4];
  if ((threadIdx.x & 63) == 0) {
    sw[threadIdx.x >> 6] = acc;
    sv[threadIdx.x >> 6] = vol;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double S = (sw[0] + sw[1]) + (sw[2] + sw[3]);
    const double V = ((sv[0] + sv[1]) + (sv[2] + sv[3])) * n2;
    *log10_out = log10(1e-16 + sqrt(S / V));
  }
}

void launch_l2_weights(const GridDims &g, const L2WeightParams &q, const double *xx0, const double *xx1, double *w2d,
                       hipStream_t st) {
  const long n = (long)g.G0 * g.G1;
  hipLaunchKernelGGL(l2_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, q, xx0, xx1, w2d);
}

void launch_l2_finalize_weights(const double *partial, long npart, const double *w2d, long nw, int n2,
                                double *log10_out, hipStream_t st) {
  hipLaunchKernelGGL(l2_finalize_weights_kernel, dim3(1), dim3(256), 0, st, partial, npart, w2d, nw, n2, log10_out);
}

long l2_partial_blocks(const GridDims &g) { return (long)((g.N0 + 63) / 64) * ((g.N1 + 3) / 4) * g.N2; }

long launch_residual_l2(const GridDims &g, const void *u, bool u_is_float, const double *aux, const double *coef,
                        const double *w2d, double *partial, const int *done, double *res_out, hipStream_t st) {
  dim3 grd((g.N0 + 63) / 64, (g.N1 + 3) / 4, g.N2), blk(256);
  if (u_is_float)
    hipLaunchKernelGGL(residual_l2_partial_kernel<float>, grd, blk, 0, st, g, (const float *)u, aux, coef, w2d,
                       partial, done, res_out);
  else
    hipLaunchKernelGGL(residual_l2_partial_kernel<double>, grd, blk, 0, st, g, (const double *)u, aux, coef, w2d,
                       partial, done, res_out);
  return (long)grd.x * grd.y * grd.z;
}

long launch_l2_partial(const GridDims &g, const double *gf, const double *w2d, double *partial, hipStream_t st) {
  dim3 grd((g.N0 + 63) / 64, (g.N1 + 3) / 4, g.N2), blk(256);
  hipLaunchKernelGGL(l2_partial_kernel, grd, blk, 0, st, g, gf, w2d, partial);
  return (long)grd.x * grd.y * grd.z;
}

void launch_l2_finalize(const double *partial, long npart, double volume, SolverDeviceState *dst, int steps_per_check,
                        double target, double *log10_out, hipStream_t st) {
  hipLaunchKernelGGL(l2_finalize_kernel, dim3(1), dim3(256), 0, st, partial, npart, volume, dst, steps_per_check,
                     target, log10_out);
}

// ---------------------------------------------------------------------------------------
// Drop-in boundary kernels.
// ---------------------------------------------------------------------------------------
// apply_bcs_inner_only.c:18-26: gfs[dst] = parity[0] * gfs[src] for both evolved gfs.
__global__ void inner_gather_kernel(long nt, int n, const int *dst, const int *src, const int8_t *par0, double *gfs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double s = (double)par0[k];
  gfs[dst[k]] = s * gfs[src[k]];
  gfs[nt + dst[k]] = s * gfs[nt + src[k]];
}

void launch_inner_gather(long nt, int n, const int *dst, const int *src, const int8_t *par0, double *gfs,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inner_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, nt, n, dst, src, par0, gfs);
}

// d f / d r at point w (0: P itself, 1: its interior neighbour) of a radiation record: the
// 11-point x0 and x1 derivatives with the record's clamped stencil shifts, combined with
// dxx0/dr and dxx1/dr (apply_bcs_outerradiation_and_inner.c:290-368).
__device__ inline double radiation_dr(const GridDims &g, const RadPoint &P, int w, const double *f) {
  const long s1 = g.G0, q = w == 0 ? P.p : P.q;
  const int o0 = P.off[w][0], o1 = P.off[w][1];
  double d0 = 0, d1 = 0;
  for (int j = 0; j < 11; j++) {
    const double c0 = o0 <= 0 ? c_fd1_shift[-o0][j] : -c_fd1_shift[o0][10 - j];
    const double c1 = o1 <= 0 ? c_fd1_shift[-o1][j] : -c_fd1_shift[o1][10 - j];
    d0 = fma(c0, f[q + (j - 5 + o0)], d0);
    d1 = fma(c1, f[q + (long)(j - 5 + o1) * s1], d1);
  }
  return P.geo[w][1] * d0 * g.invdx0 + P.geo[w][2] * d1 * g.invdx1;
}

// One pure outer point P and one gridfunction f (radiation_bcs,
// apply_bcs_outerradiation_and_inner.c:369-408). geo holds r, dxx0/dr, dxx1/dr, dxx2/dr for the
// point and for its interior neighbour; offsets are the clamped stencil shifts.
__device__ inline void radiation_point(const GridDims &g, const RadPoint &P, const double *f, double *rh, double c,
                                      double finf) {
  const double dr0 = radiation_dr(g, P, 0, f), dr1 = radiation_dr(g, P, 1, f);
  const double r = P.geo[0][0], r_int = P.geo[1][0];
  const double out_int = -c * (dr1 + (f[P.q] - finf) / r_int);
  const double kk = r_int * r_int * r_int * (rh[P.q] - out_int);
  const double rinv = 1.0 / r;
  rh[P.p] = -c * (dr0 + (f[P.p] - finf) * rinv) + kk * rinv * rinv * rinv;
}

// One (which_gz, dirn) list, one thread per (point, gridfunction).
__global__ void __launch_bounds__(64)
    radiation_list_kernel(GridDims g, int n, const RadPoint *pts, const double *gfs, double *rhs, double c_u,
        double c_v, double finf_u, double finf_v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * n) return;
  const int gf = k / n;
  radiation_point(g, pts[k % n], gfs + gf * g.ntot, rhs + gf * g.ntot, gf == 0 ? c_u : c_v,
                  gf == 0 ? finf_u : finf_v);
}

void launch_radiation_list(const GridDims &g, int n, const RadPoint *pts, const double *gfs, double *rhs, double c_u,
                           double c_v, double finf_u, double finf_v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(radiation_list_kernel, dim3((2 * n + 63) / 64), dim3(64), 0, st, g, n, pts, gfs, rhs, c_u, c_v,
                     finf_u, finf_v);
}

// All NG layers in one launch when only the x0 face has pure outer points and every layer
// lists the same (i1,i2) rows in the same order (SinhSymTP). The reference's innermost-first
// recurrence rhs_p = A_p + (r_{p-1}/r_p)^3 (rhs_{p-1} - A_{p-1}), A = -c (df/dr + (f - finf)/r),
// telescopes to rhs_p = A_p + (r_0/r_p)^3 (rhs_0 - A_0) with index 0 the last interior point, so
// every (row, gridfunction, layer) is independent: one thread each.
__global__ void __launch_bounds__(64)
    radiation_rows_kernel(GridDims g, int n, RadLayers L, const double *gfs, double *rhs, double c_u, double c_v,
        double finf_u, double finf_v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * NG * n) return;
  const int row = k % n, gf = (k / n) % 2, gz = k / (2 * n);
  const double *f = gfs + gf * g.ntot;
  double *rh = rhs + gf * g.ntot;
  const double c = gf == 0 ? c_u : c_v, finf = gf == 0 ? finf_u : finf_v;
  const RadPoint &P = L.l[gz][row], &P0 = L.l[0][row];  // P0.q: the last interior point
  const double r = P.geo[0][0], r0 = P0.geo[1][0];
  const double A = -c * (radiation_dr(g, P, 0, f) + (f[P.p] - finf) / r);
  const double A0 = -c * (radiation_dr(g, P0, 1, f) + (f[P0.q] - finf) / r0);
  const double q = r0 / r;
  rh[P.p] = A + q * q * q * (rh[P0.q] - A0);
}

void launch_radiation_rows(const GridDims &g, int n, const RadLayers &L, const double *gfs, double *rhs, double c_u,
                           double c_v, double finf_u, double finf_v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(radiation_rows_kernel, dim3((2 * NG * n + 63) / 64), dim3(64), 0, st, g, n, L, gfs, rhs, c_u,
                     c_v, finf_u, finf_v);
}

// The Einstein Toolkit thorn's outer BC (conformally_flat_BBH_apply_bcs_sommerfeld.h:16-222):
// 7-point 6th-order first derivatives, one-sided (sh = -1 backward, +1 forward, summed from the
// point outward) or centred (sh = 0, summed from +3 down to -3), then
//   rhs_p = -c (d_r f_p + (f_p - finf)/r_p) + (rhs_q + c (d_r f_q + (f_q - finf)/r_q)) (r_q/r_p)^n
// with q the neighbour one point inward; each layer reads the RHS of the layer inside it, so
// the layers are separate launches.
__device__ inline double sommerfeld_fd(const double *f, long q, long st, int sh) {
  if (sh == 0) {
    double s = (1.0 / 60.0) * f[q + 3 * st];
    s = fma(-3.0 / 20.0, f[q + 2 * st], s);
    s = fma(3.0 / 4.0, f[q + st], s);
    s = fma(-3.0 / 4.0, f[q - st], s);
    s = fma(3.0 / 20.0, f[q - 2 * st], s);
    return fma(-1.0 / 60.0, f[q - 3 * st], s);
  }
  const long d = sh * st;
  double s = (-49.0 / 20.0) * f[q];
  s = fma(6.0, f[q + d], s);
  s = fma(-15.0 / 2.0, f[q + 2 * d], s);
  s = fma(20.0 / 3.0, f[q + 3 * d], s);
  s = fma(-15.0 / 4.0, f[q + 4 * d], s);
  s = fma(6.0 / 5.0, f[q + 5 * d], s);
  s = fma(-1.0 / 6.0, f[q + 6 * d], s);
  return sh * s;
}

__device__ inline double sommerfeld_dr(const GridDims &g, const RadPoint &P, int w, const double *f) {
  const long q = w == 0 ? P.p : P.q;
  const double d0 = sommerfeld_fd(f, q, 1, P.off[w][0]) * g.invdx0;
  const double d1 = sommerfeld_fd(f, q, g.G0, P.off[w][1]) * g.invdx1;
  return P.geo[w][1] * d0 + P.geo[w][2] * d1;
}

__global__ void __launch_bounds__(64)
    sommerfeld_list_kernel(GridDims g, int n, const RadPoint *pts, const double *gfs, double *rhs,
        SommerfeldParams sp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * n) return;
  const int gf = k / n;
  const RadPoint &P = pts[k % n];
  const double *f = gfs + gf * g.ntot;
  double *rh = rhs + gf * g.ntot;
  const double c = sp.c[gf], finf = sp.finf[gf], nf = sp.falloff[gf];
  const double invr = 1.0 / P.geo[0][0];
  double val = -c * (sommerfeld_dr(g, P, 0, f) + invr * (f[P.p] - finf));
  if (nf > 0) {
    const double invr_q = 1.0 / P.geo[1][0];
    const double extrap = c * (sommerfeld_dr(g, P, 1, f) + invr_q * (f[P.q] - finf));
    val += (rh[P.q] + extrap) * pow(P.geo[1][0] * invr, nf);
  }
  rh[P.p] = val;
}

void launch_sommerfeld_list(const GridDims &g, int n, const RadPoint *pts, const double *gfs, double *rhs,
                            const SommerfeldParams &sp, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(sommerfeld_list_kernel, dim3((2 * n + 63) / 64), dim3(64), 0, st, g, n, pts, gfs, rhs, sp);
}

// ---------------------------------------------------------------------------------------
// Setup kernels (device-side setup for rfm factors, auxevol and wavespeed).
// ---------------------------------------------------------------------------------------
struct SinhParams {
  double AMAX, w, b, D;  // D = exp(1/w) - exp(-1/w)
};

// rfm_precompute_rfmstruct_define.c:8-91 (closed forms of the SinhSymTP scale factors).
__global__ void rfm_kernel(GridDims g, SinhParams s, const double *xx0, const double *xx1, rfm_struct r) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = blockIdx.y;
  if (i0 >= g.G0) return;
  const double x = xx0[i0], ep = exp(x / s.w), em = exp(-x / s.w);
  const double f0 = s.AMAX * (ep - em) / s.D;
  const double f0p = s.AMAX * (ep / s.w + em / s.w) / s.D;
  const double f0pp = s.AMAX * (ep / (s.w * s.w) - em / (s.w * s.w)) / s.D;
  const double f3sq = f0 * f0 + s.b * s.b, f3 = sqrt(f3sq);
  if (i1 == 0) {
    r.f0_of_xx0[i0] = f0;
    r.f0_of_xx0__D0[i0] = f0p;
    r.f0_of_xx0__DD00[i0] = f0pp;
    r.f0_of_xx0__DDD000[i0] = s.AMAX * (ep / (s.w * s.w * s.w) + em / (s.w * s.w * s.w)) / s.D;
    r.f3_of_xx0[i0] = f3;
    r.f3_of_xx0__D0[i0] = f0 * f0p / f3;
    r.f3_of_xx0__DD00[i0] = (f0p * f0p + f0 * f0pp) / f3 - (f0 * f0p) * (f0 * f0p) / (f3sq * f3);
  }
  const double sn = sin(xx1[i1]), cs = cos(xx1[i1]);
  if (i0 == 0) {
    r.f1_of_xx1[i1] = sn;
    r.f1_of_xx1__D1[i1] = cs;
    r.f1_of_xx1__DD11[i1] = -sn;
  }
  const double F = f0 * f0 + s.b * s.b * sn * sn, f2 = sqrt(F);
  const long k = i0 + (long)g.G0 * i1;
  r.f2_of_xx0_xx1[k] = f2;
  r.f2_of_xx0_xx1__D0[k] = f0 * f0p / f2;
  r.f2_of_xx0_xx1__D1[k] = s.b * s.b * sn * cs / f2;
  r.f2_of_xx0_xx1__DD00[k] = (f0p * f0p + f0 * f0pp) / f2 - (f0 * f0p) * (f0 * f0p) / (F * f2);
  r.f2_of_xx0_xx1__DD11[k] = s.b * s.b * (cs * cs - sn * sn) / f2 - (s.b * s.b * sn * cs) * (s.b * s.b * sn * cs) / (F * f2);
}

void launch_rfm(const GridDims &g, double AMAX, double w, double b, const double *xx0_d, const double *xx1_d,
                const rfm_struct &r, hipStream_t st) {
  SinhParams s{AMAX, w, b, exp(1.0 / w) - exp(-1.0 / w)};
  hipLaunchKernelGGL(rfm_kernel, dim3((g.G0 + 63) / 64, g.G1), dim3(64), 0, st, g, s, xx0_d, xx1_d, r);
}

// auxevol_gfs_all_points.c:9-95 at every point: psi_bg = 1 + sum m_n/(2 r_n) and the
// Bowen-York A_ij A_ij (bowen_york).
__global__ void auxevol_kernel(GridDims g, SinhParams s, Punctures pu, const double *xx0, const double *xx1,
                               const double *xx2, double *aux) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = blockIdx.y, i2 = blockIdx.z;
  if (i0 >= g.G0) return;
  const double x0 = xx0[i0], x1 = xx1[i1], x2 = xx2[i2];
  const double f0 = s.AMAX * (exp(x0 / s.w) - exp(-x0 / s.w)) / s.D;
  const double X[3] = {f0 * sin(x1) * cos(x2), f0 * sin(x1) * sin(x2), sqrt(f0 * f0 + s.b * s.b) * cos(x1)};
  double A[3][3], h[2];
  bowen_york(pu, X, A, h);
  double AA = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) AA += A[i][j] * A[i][j];
  const long p = idx3(g, i0, i1, i2);
  aux[ADD_TIMES_AUUGF * g.ntot + p] = AA;
  aux[PSI_BACKGROUNDGF * g.ntot + p] = (1.0 + h[0]) + h[1];
}

void launch_auxevol(const GridDims &g, double AMAX, double w, double b, const Punctures &pu, const double *xx0_d,
                    const double *xx1_d, const double *xx2_d, double *aux, hipStream_t st) {
  SinhParams s{AMAX, w, b, exp(1.0 / w) - exp(-1.0 / w)};
  hipLaunchKernelGGL(auxevol_kernel, dim3((g.G0 + 63) / 64, g.G1, g.G2), dim3(64), 0, st, g, s, pu, xx0_d, xx1_d,
                     xx2_d, aux);
}

// wavespeed_gf_all_points.c:9-44 from a 2-D table of min proper spacing (host-computed).
__global__ void wavespeed_kernel(GridDims g, const double *ds2d, double scale, double *aux) {
  const int i0 = NG + blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = NG + blockIdx.y, i2 = NG + blockIdx.z;
  if (i0 >= g.N0 + NG) return;
  aux[WAVESPEEDGF * g.ntot + idx3(g, i0, i1, i2)] = scale * ds2d[i0 + (long)g.G0 * i1];
}

void launch_wavespeed(const GridDims &g, const double *ds2d, double CFL, double dt, double *aux, hipStream_t st) {
  (void)dt;
  hipLaunchKernelGGL(wavespeed_kernel, dim3((g.N0 + 63) / 64, g.N1, g.N2), dim3(64), 0, st, g, ds2d, CFL, aux);
}

// Precision conversion over whole buffers (FP64 <-> FP32 shadow state).
template <typename TO, typename TI>
__global__ void convert_kernel(long n, const TI *in, TO *out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (TO)in[k];
}
void launch_convert_d2f(long n, const double *in, float *out, hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3((n + 255) / 256), dim3(256), 0, st, n, in, out);
}
void launch_convert_f2d(long n, const float *in, double *out, hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3((n + 255) / 256), dim3(256), 0, st, n, in, out);
}

// Solver tables with the wave speed folded in. The wave speed of the reference depends on
// (xx0,xx1) only (wavespeed_gf_all_points.c:19-41), so the substep kernel takes
//   c^2 K_k          (the metric coefficients, per (i0,i1)),
//   w / c^4          (L2 weights: the stage-1 stencil then gives c^2 R, R the residual),
//   0.125 c^2 AA     (per point, next to psi),
// and evaluates kv = c^2 (lap u + AA/(8 (psi+u)^7)) without reading c. A point whose wave speed
// differs from that of its (i0,i1) column on plane NG raises *bad.
// kc[6 k2 + j]: j < 5 the folded coefficients, j = 5 the folded weight, per (i0,i1) in one
// 48-byte record (three 16-byte loads in the FP64 kernel).
// Ghost phi planes (i2 < NG, i2 >= N2+NG) of the first ngf gridfunctions of y from the
// interior planes: every point takes its source's value (ghost_source; parity +1 for u and v).
// Sources lie in interior planes, so no thread reads what another writes.
template <typename T>
__global__ void fill_ghost_planes_kernel(GridDims g, T *y, int ngf) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = blockIdx.y, k = blockIdx.z;
  if (i0 >= g.G0) return;
  const int i2 = k < NG ? k : g.N2 + k;
  int s0, s1, s2;
  ghost_source(g, i0, i1, i2, s0, s1, s2);
  const long d = idx3(g, i0, i1, i2), src = idx3(g, s0, s1, s2);
  for (int f = 0; f < ngf; f++) y[f * g.ntot + d] = y[f * g.ntot + src];
}
template <typename T>
void launch_fill_ghost_planes(const GridDims &g, T *y, int ngf, hipStream_t st) {
  hipLaunchKernelGGL(fill_ghost_planes_kernel<T>, dim3((g.G0 + 63) / 64, g.G1, 2 * NG), dim3(64), 0, st, g, y, ngf);
}
template void launch_fill_ghost_planes<double>(const GridDims &, double *, int, hipStream_t);
template void launch_fill_ghost_planes<float>(const GridDims &, float *, int, hipStream_t);

// dst = src at every point, except that a point of a ghost phi plane takes src at its source
// (ghost_source): a device copy followed by fill_ghost_planes_kernel on dst, in one pass. The
// sources lie in interior planes, where the copy leaves dst equal to src.
template <typename T>
__global__ void settle_fill_kernel(GridDims g, const T *__restrict__ src, T *__restrict__ dst, int ngf) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.ntot) return;
  const int plane = g.G0 * g.G1;
  const int i2 = (int)(k / plane), r = (int)(k - (long)i2 * plane), i1 = r / g.G0, i0 = r - i1 * g.G0;
  long from = k;
  if (i2 < NG || i2 >= g.N2 + NG) {
    int s0, s1, s2;
    ghost_source(g, i0, i1, i2, s0, s1, s2);
    from = idx3(g, s0, s1, s2);
  }
  for (int f = 0; f < ngf; f++) dst[f * g.ntot + k] = src[f * g.ntot + from];
}
void launch_settle_fill(const GridDims &g, const double *src, double *dst, int ngf, hipStream_t st) {
  hipLaunchKernelGGL(settle_fill_kernel<double>, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, g, src,
                     dst, ngf);
}

__global__ void fill_kernel(long n, double v, double *out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = v;
}
void launch_fill(long n, double v, double *out, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, v, out);
}

// v += d u at every point (y = u | v, n points per gridfunction): a new damping eta + d with u and
// u_t = v - eta u unchanged
__global__ void shift_v_kernel(long n, double d, double *y) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[n + k] = fma(d, y[k], y[n + k]);
}
void launch_shift_v(long n, double d, double *y, hipStream_t st) {
  hipLaunchKernelGGL(shift_v_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, d, y);
}
