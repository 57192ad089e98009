// nrpyell_adm.hip -- the Einstein Toolkit consumer path: the converged u on the SinhSymTP grid
// to the ADMBase fields at the evolution code's Cartesian points
// (NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase.c:5-200), and the optimum damping
// parameter of the ET relaxation driver (conformally_flat_BBH_set_optimum_eta_damping.c:3-27).
//
// One thread per Cartesian point: orbital-plane permutation and shift, the closed-form inverse
// of the SinhSymTP map, a tensor-product Lagrange interpolation of u from the grid (the
// molecule placement of the Cactus AEILocalInterp operator the thorn calls), the Bowen-York
// ADM fields. u (4 MB at 128x128x16) stays in L2/MALL across the molecule reads; the 17
// outputs are written once, coalesced.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "nrpyell_common.h"
#include "nrpyell_internal.h"

namespace {

struct AdmArgs {
  int dims[3];
  double origin[3], delta[3];
  double AMAX, SINHWAA, bScale;
  Punctures etk;  // ETK-frame punctures, positions shifted
  int xy, lapse;
  double shift[3], lapse_n, eps;
  const double *u, *x, *y, *z;
  double *out[NRPYELL_ADM_NOUT];
  long n;
  int *err;  // 1: NaN coordinate after the epsilon retry, 2: point outside the grid
};

// conformally_flat_BBH_get_xx_from_xyz.c:6-20 (the closed-form inverse of the SinhSymTP map).
__device__ inline void xx_from_xyz(const AdmArgs &a, const double X[3], double xx[3]) {
  const double b = a.bScale, b2 = b * b;
  const double R2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  const double Q = sqrt(-4 * (X[2] * X[2]) * b2 + b * b * b * b + 2 * b2 * R2 + R2 * R2);
  const double ib2 = 1.0 / b2;
  const double sgn = (double)((X[2] > 0) - (X[2] < 0));
  xx[0] = a.SINHWAA * asinh(M_SQRT1_2 * sqrt(-b2 + R2 + Q) * sinh(1.0 / a.SINHWAA) / a.AMAX);
  xx[1] = acos(0.99999999 * M_SQRT1_2 * sqrt(R2 * ib2 - Q * ib2 + 1) * sgn);
  xx[2] = atan2(X[1], X[0]);
}

// Molecule of ORDER+1 points along one axis: first index and Lagrange weights. fp is x in grid
// units; the molecule spans m = MMIN..MMIN+ORDER around the nearest point (odd point count) or
// the point below x (even count), moved inward at the array edges; x more than 1e-10 grid
// spacings outside the array is an error.
template <int ORDER>
__device__ inline bool axis_weights(int n, double origin, double delta, double x, int &first, double *w) {
  const double fp = (x - origin) / delta;
  if (!(fp >= -1e-10 && fp <= (double)(n - 1) + 1e-10)) return false;
  constexpr int MMIN = -(ORDER / 2), MMAX = MMIN + ORDER;
  int c = (ORDER % 2 == 0) ? (int)floor(fp + 0.5) : (int)floor(fp);
  if (c + MMIN < 0) c = -MMIN;
  if (c + MMAX > n - 1) c = n - 1 - MMAX;
  const double t = fp - (double)c;
#pragma unroll
  for (int m = MMIN; m <= MMAX; m++) {
    double v = 1.0, den = 1.0;  // den: an integer, folded at compile time
#pragma unroll
    for (int k = MMIN; k <= MMAX; k++)
      if (k != m) {
        v *= t - (double)k;
        den *= (double)(m - k);
      }
    w[m - MMIN] = v * (1.0 / den);
  }
  first = c + MMIN;
  return true;
}

template <int ORDER>
__global__ void __launch_bounds__(256) admbase_kernel(AdmArgs a) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const double px = a.x[i], py = a.y[i], pz = a.z[i];
  const double s[3] = {px - a.shift[0], py - a.shift[1], pz - a.shift[2]};
  // (x,y,z)_NRPy = (y,z,x)_ETK for the "xy" orbital plane
  double X[3] = {a.xy ? s[1] : s[0], a.xy ? s[2] : s[1], a.xy ? s[0] : s[2]};
  double xx[3];
  xx_from_xyz(a, X, xx);
  if (isnan(xx[0] * xx[1] * xx[2])) {  // too close to a focus: move by epsilon and retry once
    X[0] += a.eps;
    X[1] += a.eps;
    X[2] += a.eps;
    xx_from_xyz(a, X, xx);
  }
  double w[3][ORDER + 1];
  int f[3];
  bool ok = !isnan(xx[0] * xx[1] * xx[2]);
  if (!ok) {
    atomicMax(a.err, 1);
  } else {
#pragma unroll
    for (int d = 0; d < 3; d++) ok = ok && axis_weights<ORDER>(a.dims[d], a.origin[d], a.delta[d], xx[d], f[d], w[d]);
    if (!ok) atomicMax(a.err, 2);
  }
  double v[NRPYELL_ADM_NOUT];
  if (ok) {
    // sum in the order k (xx2), j (xx1), i (xx0 contiguous rows)
    double acc = 0.0;
    if constexpr (ORDER >= 5) {
      // one xx2 plane of loads in flight at a time: with the loads of all (ORDER+1)^3 values
      // hoisted the kernel needs 256 VGPRs and spills (order 6: 252 of them)
#pragma unroll 1
      for (int k = 0; k <= ORDER; k++) {
        double a1 = 0.0;
        for (int j = 0; j <= ORDER; j++) {
          const double *row =
              a.u + (long)f[0] + (long)a.dims[0] * ((long)(f[1] + j) + (long)a.dims[1] * (long)(f[2] + k));
          double a0 = 0.0;
#pragma unroll
          for (int ii = 0; ii <= ORDER; ii++) a0 += w[0][ii] * row[ii];
          a1 += w[1][j] * a0;
        }
        acc += w[2][k] * a1;
      }
    } else {
      for (int k = 0; k <= ORDER; k++) {
        double a1 = 0.0;
        for (int j = 0; j <= ORDER; j++) {
          const double *row =
              a.u + (long)f[0] + (long)a.dims[0] * ((long)(f[1] + j) + (long)a.dims[1] * (long)(f[2] + k));
          double a0 = 0.0;
#pragma unroll
          for (int ii = 0; ii <= ORDER; ii++) a0 += w[0][ii] * row[ii];
          a1 += w[1][j] * a0;
        }
        acc += w[2][k] * a1;
      }
    }
    const double uu = acc;
    // conformally_flat_BBH_compute_ADM_Cartesian_quantities_from_uu.c:36-103 and the lapse
    // choices (..._lapse_power_of_psi_initial_data.c:16-35, ..._lapse_averaged_initial_data.c:18-41),
    // at the ETK point itself
    const double P3[3] = {px, py, pz};
    double A[3][3], h[2];
    bowen_york(a.etk, P3, A, h);
    const double F = h[0] + h[1];
    const double psi = ((uu + 1.0) + h[0]) + h[1];
    const double psi4 = psi * psi * psi * psi, ipsi2 = 1.0 / (psi * psi);
    v[0] = uu;
    v[1] = a.lapse ? 0.5 * (1.0 - F) / (F + 1.0) + 0.5 : pow((F + uu) + 1.0, a.lapse_n);
    v[2] = v[3] = v[4] = 0.0;
    v[5] = psi4; v[6] = 0.0; v[7] = 0.0; v[8] = psi4; v[9] = 0.0; v[10] = psi4;
    v[11] = ipsi2 * A[0][0]; v[12] = ipsi2 * A[0][1]; v[13] = ipsi2 * A[0][2];
    v[14] = ipsi2 * A[1][1]; v[15] = ipsi2 * A[1][2]; v[16] = ipsi2 * A[2][2];
  } else {
#pragma unroll
    for (int c = 0; c < NRPYELL_ADM_NOUT; c++) v[c] = __builtin_nan("");
  }
#pragma unroll
  for (int c = 0; c < NRPYELL_ADM_NOUT; c++) a.out[c][i] = v[c];
}

}  // namespace

// Drop-in for NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase (Cactus types replaced by
// plain pointers: cctk_lsh gives the point count, every array is device memory).
extern "C" int nrpyell_Initialize_ADMBase(const griddata_struct *gd, const nrpyell_admbase_options *opt,
                                          const int *cctk_lsh, const REAL *x, const REAL *y, const REAL *z,
                                          REAL *uuGF, REAL *alp, REAL *betax, REAL *betay, REAL *betaz, REAL *gxx,
                                          REAL *gxy, REAL *gxz, REAL *gyy, REAL *gyz, REAL *gzz, REAL *kxx,
                                          REAL *kxy, REAL *kxz, REAL *kyy, REAL *kyz, REAL *kzz) {
  return api_call([&]() -> int {
    if (!gd || !opt || !cctk_lsh || !x || !y || !z || !gd->gridfuncs.y_n_gfs || !gd->xx[0] || !gd->xx[1] ||
        !gd->xx[2]) {
      set_last_error("Initialize_ADMBase: null argument");
      return NRPYELL_EINVAL;
    }
    if (opt->orbital_plane != NRPYELL_ORBITAL_PLANE_XY && opt->orbital_plane != NRPYELL_ORBITAL_PLANE_XZ) {
      set_last_error("Unsupported orbital plane. Available options are \"xy\" (\"yx\") or \"xz\" (\"zx\").");
      return NRPYELL_EINVAL;
    }
    if (opt->initial_lapse != NRPYELL_LAPSE_PSI_N && opt->initial_lapse != NRPYELL_LAPSE_AVERAGED) {
      set_last_error("NRPyEllipticET does not support this lapse initial data");
      return NRPYELL_EINVAL;
    }
    if (opt->interpolation_order < 1 || opt->interpolation_order > 6) {
      set_last_error("interpolation_order must be 1..6 (Lagrange tensor product)");
      return NRPYELL_EINVAL;
    }
    REAL *outs[NRPYELL_ADM_NOUT] = {uuGF, alp, betax, betay, betaz, gxx, gxy, gxz, gyy,
                                    gyz, gzz, kxx, kxy, kxz, kyy, kyz, kzz};
    for (int c = 0; c < NRPYELL_ADM_NOUT; c++)
      if (!outs[c]) {
        set_last_error("Initialize_ADMBase: null output array");
        return NRPYELL_EINVAL;
      }
    const long n = (long)cctk_lsh[0] * cctk_lsh[1] * cctk_lsh[2];
    if (cctk_lsh[0] < 0 || cctk_lsh[1] < 0 || cctk_lsh[2] < 0) {
      set_last_error("Initialize_ADMBase: negative cctk_lsh");
      return NRPYELL_EINVAL;
    }
    if (n == 0) return NRPYELL_OK;

    const paramstruct *p = &gd->params;
    AdmArgs a;
    a.dims[0] = p->Nxx_plus_2NGHOSTS0;
    a.dims[1] = p->Nxx_plus_2NGHOSTS1;
    a.dims[2] = p->Nxx_plus_2NGHOSTS2;
    for (int d = 0; d < 3; d++) {
      if (a.dims[d] < opt->interpolation_order + 1) {
        set_last_error("Initialize_ADMBase: grid smaller than the interpolation molecule");
        return NRPYELL_EINVAL;
      }
      a.origin[d] = gd->xx[d][0];
      a.delta[d] = gd->xx[d][1] - gd->xx[d][0];
    }
    a.AMAX = p->AMAX;
    a.SINHWAA = p->SINHWAA;
    a.bScale = p->bScale;
    a.xy = opt->orbital_plane == NRPYELL_ORBITAL_PLANE_XY;
    a.lapse = opt->initial_lapse == NRPYELL_LAPSE_AVERAGED;
    a.lapse_n = opt->lapse_exponent_n;
    a.eps = opt->NRPy_epsilon;
    for (int d = 0; d < 3; d++) a.shift[d] = opt->position_shift[d];
    // The solve ran in the NRPy frame; for the "xy" plane the relaxation driver had mapped the
    // parameter file's punctures (x,y,z)_NRPy = (y,z,x)_ETK
    // (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:84-105): undo it, then shift
    // (Initialize_ADMBase.c:35-42).
    const double pos[2][3] = {{p->puncture_0_x, p->puncture_0_y, p->puncture_0_z},
                              {p->puncture_1_x, p->puncture_1_y, p->puncture_1_z}};
    const double P[2][3] = {{p->P0_x, p->P0_y, p->P0_z}, {p->P1_x, p->P1_y, p->P1_z}};
    const double S[2][3] = {{p->S0_x, p->S0_y, p->S0_z}, {p->S1_x, p->S1_y, p->S1_z}};
    const int perm[3] = {a.xy ? 2 : 0, a.xy ? 0 : 1, a.xy ? 1 : 2};  // ETK axis d = NRPy axis perm[d]
    a.etk.m[0] = p->bare_mass_0;
    a.etk.m[1] = p->bare_mass_1;
    for (int k = 0; k < 2; k++)
      for (int d = 0; d < 3; d++) {
        a.etk.pos[k][d] = pos[k][perm[d]] + a.shift[d];
        a.etk.P[k][d] = P[k][perm[d]];
        a.etk.S[k][d] = S[k][perm[d]];
      }
    a.u = gd->gridfuncs.y_n_gfs;  // gridfunction UUGF = 0
    a.x = x;
    a.y = y;
    a.z = z;
    for (int c = 0; c < NRPYELL_ADM_NOUT; c++) a.out[c] = outs[c];
    a.n = n;
    DevBuf<int> errb(1);
    int *err = errb.p;
    HIPCHK(hipMemset(err, 0, sizeof(int)));
    a.err = err;
    const dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    switch (opt->interpolation_order) {
      case 1: hipLaunchKernelGGL(admbase_kernel<1>, grd, blk, 0, 0, a); break;
      case 2: hipLaunchKernelGGL(admbase_kernel<2>, grd, blk, 0, 0, a); break;
      case 3: hipLaunchKernelGGL(admbase_kernel<3>, grd, blk, 0, 0, a); break;
      case 5: hipLaunchKernelGGL(admbase_kernel<5>, grd, blk, 0, 0, a); break;
      case 6: hipLaunchKernelGGL(admbase_kernel<6>, grd, blk, 0, 0, a); break;
      default: hipLaunchKernelGGL(admbase_kernel<4>, grd, blk, 0, 0, a); break;
    }
    HIPCHK(hipGetLastError());
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr == 1) {
      set_last_error("NAN found in (xx0,xx1,xx2) after moving the point by NRPy_epsilon; consider increasing "
                     "NRPy_epsilon");
      return NRPYELL_ENAN;
    }
    if (herr == 2) {
      set_last_error("Interpolation point outside the NRPyElliptic grid");
      return NRPYELL_EINVAL;
    }
    return NRPYELL_OK;
  });
}

// conformally_flat_BBH_set_optimum_eta_damping.c:3-16: eq. (B6) of arXiv:2111.02424.
extern "C" REAL compute_TRC(const REAL A, const REAL b, const REAL C_0, const REAL dt, const REAL dxx1,
                            const REAL dxx2) {
  const double c1 = cos(dxx1 / 2.0), c2 = cos(dxx2 / 2.0);
  const double t = A / sqrt(A * A + b * b * c1 * c1);
  return dt / (C_0 * dxx1) * atanh(t) * c1 * c2;
}

// conformally_flat_BBH_set_optimum_eta_damping.c:18-27: eta = 6.9 TRC^-0.79 - 0.7 (eq. 36).
extern "C" void set_optimum_eta_damping(const REAL dt, const REAL CFL_FACTOR, paramstruct *params) {
  const double TRC = compute_TRC(params->AMAX, params->bScale, CFL_FACTOR, dt, params->dxx1, params->dxx2);
  params->eta_damping = 6.9 * pow(TRC, -0.79) - 0.7;
}
