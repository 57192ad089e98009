/*
 * oracle.c -- CPU restatement of NRPyElliptic's hyperbolic-relaxation hot path.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker. The product
 * library never calls it.
 *
 * Written from the reference's behaviour (the NRPyElliptic_Ccodes sources), not copied:
 * each function cites the reference file:line it restates.  Where the reference
 * rounding is fixed by explicit FMA intrinsics (rhs_eval, residual, MoL stage
 * updates) the same FMA tree is spelled out with fma(), so these are bit-exact.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_set_grid           bit-exact vs the compiled reference (oracle/_ref)
 *   orc_rfm_define         <=1e-14 rel vs oracle/_ref; closed forms f2^2=f0^2+b^2 s^2, f3^2=f0^2+b^2
 *   orc_auxevol            <=1e-12 rel vs oracle/_ref; psi closed form; AA^2 vs central
 *                          differences of the Bowen-York vector field V
 *   orc_find_timestep      <=1e-15 rel vs oracle/_ref
 *   orc_wavespeed          <=1e-14 rel vs oracle/_ref
 *   orc_bcstruct           bit-exact (dst, src, 10 parities, outer lists) vs oracle/_ref;
 *                          invariants: every src is interior or pure outer
 *   orc_rhs_eval           bit-exact vs oracle/_ref; 10th-order convergence on analytic u
 *   orc_residual           bit-exact vs oracle/_ref
 *   orc_apply_bcs_*        <=1e-13 rel vs oracle/_ref (geometry restated from the
 *                          SinhSymTP map, not from the reference's CSE)
 *   orc_mol_step           <=1e-13 rel vs oracle/_ref after 1 and 20 steps
 *   orc_l2_norm            <=1e-13 abs on log10 vs oracle/_ref (1 OpenMP thread)
 *
 * Einstein Toolkit consumer path (NRPyEllipticET; its sources need the Cactus headers, so
 * they cannot be compiled here, tests/test_oracle_adm.py):
 *   orc_get_xx_from_xyz    vs the compiled reference's SymTP inverse map
 *                          Cart_to_xx_and_nearest_i0i1i2 (sqrt arguments bit-identical):
 *                          xx0 = w asinh(xx0_SymTP sinh(1/w)/AMAX), cos xx1 = 0.99999999 cos xx1_SymTP
 *                          to 1e-15; round trip through the SinhSymTP map
 *   orc_lagrange_interp3   exact for tensor polynomials of degree <= order; molecule placement
 *                          of the Cactus AEILocalInterp "Lagrange polynomial interpolation"
 *                          operator (not in the reference tree: placement parity unpinned)
 *   orc_adm_point          psi^12 K_ij K_ij = the compiled reference's auxevol AA^2 at grid
 *                          points (<=1e-12 rel); K_ij traceless; surface integrals
 *                          (1/8pi) \oint K_ij n^j dA = P_i and (1/8pi) \oint eps_ijk x^j K^kl n_l dA = S_i
 *                          around a single puncture; lapse closed forms
 *   orc_compute_TRC        atanh written as a log; eta(TRC = 1) = 6.2
 *   orc_apply_bcs_sommerfeld  the thorn's outer BC: 7-point weights regenerated as Fornberg
 *                          rationals; exact on polynomials of degree <= 6; d_r r = 1 and
 *                          d_r of the Cartesian z equal to z/r (Jacobian orientation);
 *                          f = k/r gives rhs = 0 and an interior rhs K/r^3 is carried to
 *                          K/r^3 at every outer layer (falloff power 3)
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NG ORC_NGHOSTS
#define IDX3(g, i0, i1, i2) ((i0) + (g)->Ng[0] * ((i1) + (g)->Ng[1] * (i2)))
#define NTOT(g) ((long)(g)->Ng[0] * (g)->Ng[1] * (g)->Ng[2])

/* ---------------------------------------------------------------------------------
 * Grid coordinates. set_Nxx_dxx_invdx_params__and__xx.c:21-53: cell-centred
 * uniform grids; EigenCoord 0 -> xx0 in [0,1], 1 -> xx0 in [0,AMAX]; xx1 in [0,pi];
 * xx2 in [-pi,pi]. xx[j] = xxmin + ((j-NGHOSTS)+1/2)*dxx (the reference's GCC build
 * contracts this to one fma, which is what makes it bit-exact here).
 * ------------------------------------------------------------------------------- */
int orc_set_grid(orc_grid *g, int eigen, double *xx0, double *xx1, double *xx2) {
  const double pi = 3.141592653589793238462643383279502884;
  double lo[3] = {0.0, 0.0, -pi}, hi[3] = {eigen ? g->AMAX : 1.0, pi, pi};
  double *xx[3] = {xx0, xx1, xx2};
  for (int d = 0; d < 3; d++) {
    if (g->N[d] < 1) return -1;
    g->Ng[d] = g->N[d] + 2 * NG;
    g->dxx[d] = (hi[d] - lo[d]) / (double)g->N[d];
    g->invdx[d] = 1.0 / g->dxx[d];
    for (int j = 0; j < g->Ng[d]; j++) xx[d][j] = fma((double)(j - NG) + 0.5, g->dxx[d], lo[d]);
  }
  return 0;
}

/* SinhSymTP radial map pieces: f0 = AMAX sinh(x/w)/sinh(1/w) written with exp()
 * exactly as the reference evaluates it (rfm_precompute_rfmstruct_define.c:10). */
static double sinh_den(const orc_grid *g) { return exp(1.0 / g->SINHWAA) - exp(-1 / g->SINHWAA); }

/* rfm_precompute_rfmstruct_define.c:8-91. Closed forms of the SinhSymTP scale
 * factors and their derivatives (reference_metric.py:308-400):
 *   f0 = AMAX sinh(x0/w)/sinh(1/w), f1 = sin x1, f2 = sqrt(f0^2 + b^2 sin^2 x1),
 *   f3 = sqrt(f0^2 + b^2). */
void orc_rfm_define(const orc_grid *g, const double *xx0, const double *xx1, orc_rfm *rfm) {
  const double w = g->SINHWAA, A = g->AMAX, b = g->bScale, D = sinh_den(g);
  const int N0 = g->Ng[0], N1 = g->Ng[1];
  for (int i = 0; i < N0; i++) {
    const double x = xx0[i], ep = exp(x / w), em = exp(-x / w);
    const double f0 = A * (ep - em) / D;
    const double f0p = A * (ep / w + em / w) / D;
    const double f0pp = A * (ep / (w * w) - em / (w * w)) / D;
    rfm->f0[i] = f0;
    rfm->f0_D0[i] = f0p;
    rfm->f0_DD00[i] = f0pp;
    rfm->f0_DDD000[i] = A * (ep / (w * w * w) + em / (w * w * w)) / D;
    const double f3 = sqrt(f0 * f0 + b * b);
    rfm->f3[i] = f3;
    rfm->f3_D0[i] = f0 * f0p / f3;
    rfm->f3_DD00[i] = (f0p * f0p + f0 * f0pp) / f3 - (f0 * f0p) * (f0 * f0p) / (f3 * f3 * f3);
  }
  for (int j = 0; j < N1; j++) {
    rfm->f1[j] = sin(xx1[j]);
    rfm->f1_D1[j] = cos(xx1[j]);
    rfm->f1_DD11[j] = -sin(xx1[j]);
  }
  for (int j = 0; j < N1; j++)
    for (int i = 0; i < N0; i++) {
      const double f0 = rfm->f0[i], f0p = rfm->f0_D0[i], f0pp = rfm->f0_DD00[i];
      const double s = sin(xx1[j]), c = cos(xx1[j]);
      const double F = f0 * f0 + b * b * s * s, f2 = sqrt(F);
      const long k = i + (long)N0 * j;
      rfm->f2[k] = f2;
      rfm->f2_D0[k] = f0 * f0p / f2;
      rfm->f2_D1[k] = b * b * s * c / f2;
      rfm->f2_DD00[k] = (f0p * f0p + f0 * f0pp) / f2 - (f0 * f0p) * (f0 * f0p) / (F * f2);
      rfm->f2_DD11[k] = b * b * (c * c - s * s) / f2 - (b * b * s * c) * (b * b * s * c) / (F * f2);
    }
}

/* Cartesian position of a SinhSymTP point (xx_to_Cart.c; reference_metric.py:308-400). */
static void sinhsymtp_to_cart(const orc_grid *g, double D, double x0, double x1, double x2, double X[3]) {
  const double f0 = g->AMAX * (exp(x0 / g->SINHWAA) - exp(-x0 / g->SINHWAA)) / D;
  X[0] = f0 * sin(x1) * cos(x2);
  X[1] = f0 * sin(x1) * sin(x2);
  X[2] = sqrt(f0 * f0 + g->bScale * g->bScale) * cos(x1);
}

/* auxevol_gfs_all_points.c:9-95 at ALL points (ghosts included):
 *   psi_bg = 1 + sum_n m_n/(2 r_n)                       (NRPyElliptic_SourceTerm.py:191-203)
 *   AA2    = A_ij A_ij,  A_ij = sum_n [d_j V_i + d_i V_j - (2/3) delta_ij d_k V_k],
 *   V^i    = -(7/4) P^i/r - (1/4)(x.P) x^i/r^3 + (x cross S)^i/r^3 (SourceTerm.py:59-70,121-142)
 * Differentiating V by hand gives, per puncture with x relative to it:
 *   A_ij = (3/2)(P_i x_j + P_j x_i)/r^3 - (3/2)(x.P) delta_ij/r^3 + (3/2)(x.P) x_i x_j/r^5
 *          - 3[(x cross S)_i x_j + (x cross S)_j x_i]/r^5                                   */
void orc_auxevol(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2, double *aux) {
  const double D = sinh_den(g);
  const long ntot = NTOT(g);
#pragma omp parallel for
  for (int i2 = 0; i2 < g->Ng[2]; i2++)
    for (int i1 = 0; i1 < g->Ng[1]; i1++)
      for (int i0 = 0; i0 < g->Ng[0]; i0++) {
        double X[3];
        sinhsymtp_to_cart(g, D, xx0[i0], xx1[i1], xx2[i2], X);
        double A[3][3] = {{0}}, psi = 1.0;
        for (int n = 0; n < 2; n++) {
          const double x[3] = {X[0] - g->pos[n][0], X[1] - g->pos[n][1], X[2] - g->pos[n][2]};
          const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], r = sqrt(r2);
          const double ir3 = 1.0 / (r2 * r), ir5 = ir3 / r2;
          const double *P = g->P[n], *S = g->S[n];
          const double xP = x[0] * P[0] + x[1] * P[1] + x[2] * P[2];
          const double xS[3] = {x[1] * S[2] - x[2] * S[1], x[2] * S[0] - x[0] * S[2], x[0] * S[1] - x[1] * S[0]};
          psi += g->mass[n] / (2.0 * r);
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
              A[i][j] += 1.5 * (P[i] * x[j] + P[j] * x[i]) * ir3 - (i == j ? 1.5 * xP * ir3 : 0.0) +
                         1.5 * xP * x[i] * x[j] * ir5 - 3.0 * (xS[i] * x[j] + xS[j] * x[i]) * ir5;
        }
        double AA = 0.0;
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) AA += A[i][j] * A[i][j];
        const long p = IDX3(g, i0, i1, i2);
        aux[ORC_ADD_TIMES_AUU * ntot + p] = AA;
        aux[ORC_PSI_BACKGROUND * ntot + p] = psi;
      }
}

/* Proper grid spacings of SinhSymTP (find_timestep.c:19-32):
 *   ds0 = dxx0 f0' f2/f3, ds1 = dxx1 f2, ds2 = dxx2 f0 sin x1. */
static void proper_ds(const orc_grid *g, double D, double x0, double x1, double ds[3]) {
  const double iw = 1.0 / g->SINHWAA, ep = exp(iw * x0), em = exp(-iw * x0);
  const double s = sin(x1), f0sq = g->AMAX * g->AMAX * (ep - em) * (ep - em) / (D * D);
  const double f2 = sqrt(g->bScale * g->bScale * s * s + f0sq), AD = g->AMAX / D;
  ds[0] = g->dxx[0] * AD * f2 * (iw * ep + iw * em) / sqrt(g->bScale * g->bScale + f0sq);
  ds[1] = g->dxx[1] * f2;
  ds[2] = g->dxx[2] * s * AD * (ep - em);
}

/* find_timestep.c:8-45: dt = min over interior of min(|ds0|,|ds1|,|ds2|); the CFL
 * factor is NOT applied (find_timestep.c:45 returns dsmin*1.0/1.0). */
double orc_find_timestep(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2) {
  const double D = sinh_den(g);
  double dsmin = 1e38;
  for (int i2 = NG; i2 < NG + g->N[2]; i2++)
    for (int i1 = NG; i1 < NG + g->N[1]; i1++)
      for (int i0 = NG; i0 < NG + g->N[0]; i0++) {
        double ds[3];
        proper_ds(g, D, xx0[i0], xx1[i1], ds);
        const double m = fmin(fabs(ds[0]), fmin(fabs(ds[1]), fabs(ds[2])));
        if (m < dsmin) dsmin = m;
      }
  return dsmin;
}

/* wavespeed_gf_all_points.c:9-44: c = CFL*min(ds0,ds1,ds2)/dt at interior points. */
void orc_wavespeed(const orc_grid *g, double CFL, double dt, const double *xx0, const double *xx1,
                   const double *xx2, double *aux) {
  const double D = sinh_den(g);
  const long ntot = NTOT(g);
  for (int i2 = NG; i2 < NG + g->N[2]; i2++)
    for (int i1 = NG; i1 < NG + g->N[1]; i1++)
      for (int i0 = NG; i0 < NG + g->N[0]; i0++) {
        double ds[3];
        proper_ds(g, D, xx0[i0], xx1[i1], ds);
        const double m = ds[0] < (ds[1] < ds[2] ? ds[1] : ds[2]) ? ds[0] : (ds[1] < ds[2] ? ds[1] : ds[2]);
        aux[ORC_WAVESPEED * ntot + IDX3(g, i0, i1, i2)] = CFL * m / dt;
      }
}

/* compute_wavespeed_at_OB.c:10: wavespeed at (Ng0-NGHOSTS-1, NGHOSTS, Ng2/2). */
double orc_wavespeed_at_OB(const orc_grid *g, const double *aux) {
  return aux[ORC_WAVESPEED * NTOT(g) + IDX3(g, g->Ng[0] - NG - 1, NG, g->Ng[2] / 2)];
}

/* ---------------------------------------------------------------------------------
 * Ghost-zone maps. bcstruct_set_up.c:27-194 (eigen-coordinate round trip),
 * :204-288 (parities), :350-521 (tables).
 * ------------------------------------------------------------------------------- */

/* Round trip of one point through the SymTP eigen-coordinate Cartesian map
 * (bcstruct_set_up.c:57-113). Returns 0, or -1 where the reference would abort
 * with "Cartesian disagreement" (:173-185). */
static int eigen_round_trip(const orc_grid *g, const double *xx[3], int i0, int i1, int i2, int out[3]) {
  const double b = g->bScale, b2 = b * b;
  const double x0 = xx[0][i0], x1 = xx[1][i1], x2 = xx[2][i2];
  /* SymTP (the eigen-coordinate): x = x0 sin x1 cos x2, y = x0 sin x1 sin x2, z = sqrt(b^2+x0^2) cos x1 */
  const double X = x0 * sin(x1) * cos(x2), Y = x0 * sin(x1) * sin(x2), Z = sqrt(b2 + x0 * x0) * cos(x1);
  /* Inverse SymTP map (prolate spheroidal inversion). */
  const double R2 = X * X + Y * Y + Z * Z;
  const double Q = sqrt(-4 * Z * Z * b2 + b2 * b2 + 2 * b2 * R2 + R2 * R2);
  const double sgnZ = (double)((Z > 0) - (Z < 0));
  const double y[3] = {sqrt(0.5) * sqrt(-b2 + R2 + Q), acos(sqrt(0.5) * sqrt(R2 / b2 - Q / b2 + 1) * sgnZ),
                       atan2(Y, X)};
  for (int d = 0; d < 3; d++) {
    const double dx = g->dxx[d];
    const double xmin = xx[d][0] - ((double)(0 - NG) + 0.5) * dx;
    out[d] = (int)((y[d] - xmin - 0.5 * dx + (double)NG * dx) / dx + 0.5);
  }
  /* Sanity check in the ORIGINAL (SinhSymTP) Cartesian map, relative tolerance 1e-8. */
  const double D = sinh_den(g);
  double A[3], B[3];
  sinhsymtp_to_cart(g, D, x0, x1, x2, A);
  sinhsymtp_to_cart(g, D, xx[0][out[0]], xx[1][out[1]], xx[2][out[2]], B);
  const double nrm = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) + 1e-15;
  for (int d = 0; d < 3; d++)
    if (fabs(A[d] - B[d]) > 1e-8 * nrm) return -1;
  return 0;
}

/* Parity of the 10 NRPy+ tensor types between a ghost point and its image: dot products
 * of the SinhSymTP unit basis vectors at both points (bcstruct_set_up.c:221-286).
 * e0 ~ (f3 s cos x2, f3 s sin x2, f0 c)/f2, e1 ~ (f0 c cos x2, f0 c sin x2, -f3 s)/f2,
 * e2 = (-sin x2, cos x2, 0). Evaluated, as the reference does, on the coordinates handed
 * to bcstruct_set_up (eigen grid) with the SinhSymTP scale factors: where exp()
 * overflows the dot products are NaN and the sign test leaves +1 (:283-286). */
static void parities(const orc_grid *g, const double p[3], const double q[3], int8_t par[10]) {
  const double iw = 1.0 / g->SINHWAA, b2 = g->bScale * g->bScale;
  const double Dd = exp(iw) - exp(-iw), K = g->AMAX * g->AMAX / (Dd * Dd);
  const double sp = exp(iw * p[0]) - exp(-iw * p[0]), sq = exp(iw * q[0]) - exp(-iw * q[0]);
  const double s1p = sin(p[1]), s1q = sin(q[1]);
  const double f0p2 = sp * sp * K, f0q2 = sq * sq * K;
  const double inv_f2f2 = 1 / (sqrt(f0q2 + b2 * s1q * s1q) * sqrt(b2 * s1p * s1p + f0p2));
  const double f0f0cc = inv_f2f2 * sp * sq * K * cos(p[1]) * cos(q[1]);
  const double f3f3ss = inv_f2f2 * s1p * s1q * sqrt(f0q2 + b2) * sqrt(b2 + f0p2);
  const double ss = sin(p[2]) * sin(q[2]), cc = cos(p[2]) * cos(q[2]);
  const double d0 = f0f0cc + ss * f3f3ss + f3f3ss * cc;
  const double d1 = f0f0cc * ss + f0f0cc * cc + f3f3ss;
  const double d2 = ss + cc;
  const double v[10] = {1, d0, d1, d2, d0 * d0, d0 * d1, d0 * d2, d1 * d1, d1 * d2, d2 * d2};
  for (int k = 0; k < 10; k++) par[k] = (v[k] < 0) ? -1 : 1;
}

static int is_interior(const orc_grid *g, int i0, int i1, int i2) {
  return i0 >= NG && i0 < g->Ng[0] - NG && i1 >= NG && i1 < g->Ng[1] - NG && i2 >= NG && i2 < g->Ng[2] - NG;
}

/* bcstruct_set_up.c:350-521. Returns 0, or -1 on a round-trip failure (the reference exits). */
int orc_bcstruct(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2, orc_bc *bc) {
  const double *xx[3] = {xx0, xx1, xx2};
  int n_inner = 0;
  for (int i2 = 0; i2 < g->Ng[2]; i2++)
    for (int i1 = 0; i1 < g->Ng[1]; i1++)
      for (int i0 = 0; i0 < g->Ng[0]; i0++) {
        if (is_interior(g, i0, i1, i2)) continue;
        int t[3];
        if (eigen_round_trip(g, xx, i0, i1, i2, t)) return -1;
        if (t[0] == i0 && t[1] == i1 && t[2] == i2) continue; /* pure outer */
        orc_inner *e = &bc->inner[n_inner++];
        e->dstpt = IDX3(g, i0, i1, i2);
        e->srcpt = IDX3(g, t[0], t[1], t[2]);
        const double p[3] = {xx0[i0], xx1[i1], xx2[i2]}, q[3] = {xx0[t[0]], xx1[t[1]], xx2[t[2]]};
        parities(g, p, q, e->parity);
      }
  bc->num_inner = n_inner;

  /* Face ranges grow outward one layer at a time (bcstruct_set_up.c:411-467). */
  int lo[3] = {NG, NG, NG}, hi[3] = {g->Ng[0] - NG, g->Ng[1] - NG, g->Ng[2] - NG};
  int range[NG][6][6];
  for (int gz = 0; gz < NG; gz++) {
    for (int d = 0; d < 3; d++) {
      int *rl = range[gz][2 * d], *rh = range[gz][2 * d + 1];
      for (int e = 0; e < 3; e++) {
        rl[2 * e] = lo[e]; rl[2 * e + 1] = hi[e];
      }
      rl[2 * d] = lo[d] - 1; rl[2 * d + 1] = lo[d];
      lo[d]--;
      for (int e = 0; e < 3; e++) {
        rh[2 * e] = lo[e]; rh[2 * e + 1] = hi[e];
      }
      rh[2 * d] = hi[d]; rh[2 * d + 1] = hi[d] + 1;
      hi[d]++;
    }
  }
  for (int gz = 0; gz < NG; gz++)
    for (int dirn = 0; dirn < 3; dirn++) {
      int n = 0;
      for (int side = 0; side < 2; side++) {
        const int face = 2 * dirn + side;
        const int *r = range[gz][face];
        for (int i2 = r[4]; i2 < r[5]; i2++)
          for (int i1 = r[2]; i1 < r[3]; i1++)
            for (int i0 = r[0]; i0 < r[1]; i0++) {
              int t[3];
              if (eigen_round_trip(g, xx, i0, i1, i2, t)) return -1;
              if (!(t[0] == i0 && t[1] == i1 && t[2] == i2)) continue;
              orc_outer *o = &bc->outer[dirn + 3 * gz][n++];
              o->i0 = (short)i0; o->i1 = (short)i1; o->i2 = (short)i2;
              o->FACEX0 = (int8_t)((face == 0) - (face == 1));
              o->FACEX1 = (int8_t)((face == 2) - (face == 3));
              o->FACEX2 = (int8_t)((face == 4) - (face == 5));
            }
      }
      bc->num_outer[dirn + 3 * gz] = n;
    }
  return 0;
}

/* ---------------------------------------------------------------------------------
 * Interior operator. rhs_eval.c:7-117 and residual_all_points.c:7-112.
 * 10th-order centred FD (weights rhs_eval.c:23-27; Fornberg):
 *   d1 taps k=1..5: 5/6, -5/21, 5/84, -5/504, 1/1260
 *   d2 centre -5269/1800, taps k=1..5: 5/3, -5/21, 5/126, -5/1008, 1/3150
 * The FMA tree below is the reference's (SIMD_intrinsics.h:18-21 semantics).
 * ------------------------------------------------------------------------------- */
typedef struct { double dD0, dD1, dDD00, dDD11, dDD22; } fdset_t;

static inline double fd1(const double m[6], const double p[6]) {
  /* m[k] = u(i-k), p[k] = u(i+k) */
  return fma(5.0 / 504.0, m[4] - p[4],
             fma(5.0 / 6.0, p[1] - m[1],
                 fma(5.0 / 84.0, p[3] - m[3], fma(1.0 / 1260.0, p[5] - m[5], (5.0 / 21.0) * (m[2] - p[2])))));
}
static inline double fd2(const double m[6], const double p[6], double c0) {
  /* c0 = (5269/1800) u(i) */
  return fma(5.0 / 126.0, m[3] + p[3],
             fma(5.0 / 3.0, m[1] + p[1],
                 fma(1.0 / 3150.0, m[5] + p[5], -fma(5.0 / 1008.0, m[4] + p[4], fma(5.0 / 21.0, m[2] + p[2], c0)))));
}

static fdset_t fd_all(const orc_grid *g, const double *u, int i0, int i1, int i2) {
  double m0[6], p0[6], m1[6], p1[6], m2[6], p2[6];
  for (int k = 1; k <= 5; k++) {
    m0[k] = u[IDX3(g, i0 - k, i1, i2)]; p0[k] = u[IDX3(g, i0 + k, i1, i2)];
    m1[k] = u[IDX3(g, i0, i1 - k, i2)]; p1[k] = u[IDX3(g, i0, i1 + k, i2)];
    m2[k] = u[IDX3(g, i0, i1, i2 - k)]; p2[k] = u[IDX3(g, i0, i1, i2 + k)];
  }
  const double c0 = (5269.0 / 1800.0) * u[IDX3(g, i0, i1, i2)];
  fdset_t f;
  f.dD0 = g->invdx[0] * fd1(m0, p0);
  f.dD1 = g->invdx[1] * fd1(m1, p1);
  f.dDD00 = (g->invdx[0] * g->invdx[0]) * fd2(m0, p0, c0);
  f.dDD11 = (g->invdx[1] * g->invdx[1]) * fd2(m1, p1, c0);
  f.dDD22 = (g->invdx[2] * g->invdx[2]) * fd2(m2, p2, c0);
  return f;
}

/* flat-space Laplacian of u in SinhSymTP plus the Bowen-York source term
 * AA2/(8 (psi+u)^7)  (NRPyElliptic_RHSs.py:104-117; rhs_eval.c:98-110 expression tree) */
static double laplacian_plus_source(const orc_grid *g, const orc_rfm *r, fdset_t f, double uu, double psi,
                                    double AA, int i0, int i1) {
  const long k2 = i0 + (long)g->Ng[0] * i1;
  const double f0 = r->f0[i0], f0p = r->f0_D0[i0], f0pp = r->f0_DD00[i0];
  const double f1 = r->f1[i1], f1p = r->f1_D1[i1];
  const double f2 = r->f2[k2], f2_0 = r->f2_D0[k2];
  const double f3 = r->f3[i0], f3p = r->f3_D0[i0];
  const double a0 = 1.0 / (f2 * f2);                  /* 1/f2^2         */
  const double a2 = (f3 * f3) / (f0p * f0p);          /* f3^2/f0'^2     */
  const double a3 = 1.0 / ((f2 * f2) * f2);           /* 1/f2^3         */
  const double X = psi + uu;
  const double X7 = (((((X * X) * X) * X) * X) * X) * X;
  const double num = fma(f3, fma(f0p, f2_0, f0pp * f2), -(f0p * (f2 * f3p)));
  const double tA = num / ((f0p * f0p) * f0p);
  const double tB = ((a0 * -1.0) * ((f3 * f3) / f0p)) / f0;
  const double inner = fma(a2, a3 * f2_0, -tB);
  const double B0 = fma(a3 * f3, tA, -inner);
  double e = fma(a0, f.dDD11, -(f.dD0 * B0));
  e = fma(a0 * f1p, f.dD1 / f1, e);
  e = fma(1.0 / 8.0, AA / X7, e);
  e = fma(a0, a2 * f.dDD00, e);
  e = fma(f.dDD22, (1.0 / (f1 * f1)) / (f0 * f0), e);
  return e;
}

void orc_rhs_eval(const orc_grid *g, const orc_rfm *rfm, const double *aux, const double *in, double *rhs) {
  const long nt = NTOT(g);
  const double *u = in, *v = in + nt;
#pragma omp parallel for
  for (int i2 = NG; i2 < NG + g->N[2]; i2++)
    for (int i1 = NG; i1 < NG + g->N[1]; i1++)
      for (int i0 = NG; i0 < NG + g->N[0]; i0++) {
        const long p = IDX3(g, i0, i1, i2);
        const fdset_t f = fd_all(g, u, i0, i1, i2);
        const double c = aux[ORC_WAVESPEED * nt + p];
        const double e = laplacian_plus_source(g, rfm, f, u[p], aux[ORC_PSI_BACKGROUND * nt + p],
                                               aux[ORC_ADD_TIMES_AUU * nt + p], i0, i1);
        rhs[p] = fma(-g->eta, u[p], v[p]); /* NegFusedMulAdd: v - eta*u, one rounding */
        rhs[nt + p] = (c * c) * e;
      }
}

void orc_residual(const orc_grid *g, const orc_rfm *rfm, const double *aux, const double *in, double *res) {
  const long nt = NTOT(g);
#pragma omp parallel for
  for (int i2 = NG; i2 < NG + g->N[2]; i2++)
    for (int i1 = NG; i1 < NG + g->N[1]; i1++)
      for (int i0 = NG; i0 < NG + g->N[0]; i0++) {
        const long p = IDX3(g, i0, i1, i2);
        const fdset_t f = fd_all(g, in, i0, i1, i2);
        res[p] = laplacian_plus_source(g, rfm, f, in[p], aux[ORC_PSI_BACKGROUND * nt + p],
                                       aux[ORC_ADD_TIMES_AUU * nt + p], i0, i1);
      }
}

/* ---------------------------------------------------------------------------------
 * Boundary conditions. apply_bcs_inner_only.c:11-27; apply_bcs_outerradiation_and_inner.c.
 * ------------------------------------------------------------------------------- */
void orc_apply_bcs_inner_only(const orc_grid *g, const orc_bc *bc, double *gfs) {
  const long nt = NTOT(g);
  for (int gf = 0; gf < 2; gf++)
    for (int k = 0; k < bc->num_inner; k++) {
      const orc_inner *e = &bc->inner[k];
      /* evol_gf_parity = {0,0} (NRPy_basic_defines.h:199): scalar parity */
      gfs[gf * nt + e->dstpt] = e->parity[0] * gfs[gf * nt + e->srcpt];
    }
}

/* 11-point first-derivative weights, stencil i0+offset-5 .. i0+offset+5 (Fornberg;
 * the same rationals as apply_bcs_outerradiation_and_inner.c:11-141). */
static const int FD1_NUM[11][11] = {
  {1, -10, 45, -120, 35, -252, 105, -40, 45, -10, 7381},    /* offset -5 */
  {-1, 1, -9, 2, -21, 63, -7, 6, -9, 4609, 1},              /* -4 */
  {1, -2, 1, -8, 7, -28, 7, -8, 341, 2, -1},                /* -3 */
  {-1, 1, -3, 1, -7, 21, -7, 319, 3, -1, 1},                /* -2 */
  {1, -1, 3, -4, 1, -6, 11, 4, -3, 1, -1},                  /* -1 */
  {-1, 5, -5, 5, -5, 0, 5, -5, 5, -5, 1},                   /*  0 */
  {1, -1, 3, -4, -11, 6, -1, 4, -3, 1, -1},                 /*  1 */
  {-1, 1, -3, -319, 7, -21, 7, -1, 3, -1, 1},               /*  2 */
  {1, -2, -341, 8, -7, 28, -7, 8, -1, 2, -1},               /*  3 */
  {-1, -4609, 9, -6, 7, -63, 21, -2, 9, -1, 1},             /*  4 */
  {-7381, 10, -45, 40, -105, 252, -35, 120, -45, 10, -1}};  /*  5 */
static const int FD1_DEN[11][11] = {
  {10, 9, 8, 7, 1, 5, 2, 1, 2, 1, 2520},
  {90, 8, 14, 1, 5, 10, 1, 1, 2, 2520, 10},
  {360, 63, 6, 15, 6, 15, 3, 3, 280, 9, 90},
  {840, 72, 40, 4, 12, 20, 4, 420, 8, 24, 360},
  {1260, 105, 56, 21, 2, 5, 30, 7, 28, 63, 840},
  {1260, 504, 84, 21, 6, 1, 6, 21, 84, 504, 1260},
  {840, 63, 28, 7, 30, 5, 2, 21, 56, 105, 1260},
  {360, 24, 8, 420, 4, 20, 12, 4, 40, 72, 840},
  {90, 9, 280, 3, 3, 15, 6, 15, 6, 63, 360},
  {10, 2520, 2, 1, 1, 10, 5, 1, 14, 8, 90},
  {2520, 1, 2, 1, 2, 5, 1, 7, 8, 9, 10}};

double orc_fd1_weight(int offset, int j) {
  return (double)FD1_NUM[offset + 5][j] / (double)FD1_DEN[offset + 5][j];
}

/* First derivative along x0 (dir 0) or x1 (dir 1) with a stencil shifted by `offset`
 * (FD1_arbitrary_upwind_x{0,1}_dirn, apply_bcs_outerradiation_and_inner.c:6-286).
 * Terms are summed from the lowest tap upward, skipping a zero centre weight. */
static double fd1_shifted(const orc_grid *g, const double *f, int dir, int i0, int i1, int i2, int offset) {
  double s = 0.0;
  int first = 1;
  for (int j = 0; j < 11; j++) {
    if (FD1_NUM[offset + 5][j] == 0) continue;
    const int d = offset - 5 + j;
    const double val = dir == 0 ? f[IDX3(g, i0 + d, i1, i2)] : f[IDX3(g, i0, i1 + d, i2)];
    s = first ? orc_fd1_weight(offset, j) * val : fma(orc_fd1_weight(offset, j), val, s);
    first = 0;
  }
  return s * g->invdx[dir];
}

/* r and d(xx0)/dr, d(xx1)/dr at fixed spherical angles for SinhSymTP
 * (r_and_partial_xi_partial_r_derivs, apply_bcs_outerradiation_and_inner.c:290-321),
 * restated from the map: r^2 = f0^2 + b^2 cos^2 x1, cos(theta) = f3 cos x1 / r.
 * With g0 = d(cos theta)/dxx0, g1 = d(cos theta)/dxx1 and J00 = dr/dxx0, J01 = dr/dxx1,
 * inverting the 2x2 Jacobian gives dxx0/dr = g1/det, dxx1/dr = -g0/det,
 * det = J00 g1 - J01 g0. */
static void radial_geometry(const orc_grid *g, double x0, double x1, double out[3]) {
  const double w = g->SINHWAA, A = g->AMAX, b = g->bScale, D = sinh_den(g);
  const double ep = exp(x0 / w), em = exp(-x0 / w);
  const double f0 = A * (ep - em) / D, f0p = A * (ep + em) / (w * D);
  const double f3 = sqrt(f0 * f0 + b * b), f3p = f0 * f0p / f3;
  const double s = sin(x1), c = cos(x1);
  const double r = sqrt(f0 * f0 + b * b * c * c);
  const double J00 = f0 * f0p / r, J01 = -b * b * s * c / r;
  const double g0 = f3p * c / r - f3 * c * J00 / (r * r);
  const double g1 = -f3 * s / r - f3 * c * J01 / (r * r);
  const double det = J00 * g1 - J01 * g0;
  out[0] = r;
  out[1] = g1 / det;
  out[2] = -g0 / det;
}

/* d_r f at a point: stencil offset away from the face and clamped in-bounds
 * (compute_partial_r_f, apply_bcs_outerradiation_and_inner.c:325-364). */
static double partial_r_f(const orc_grid *g, const double *f, int i0, int i1, int i2, const int face[3],
                          const double geo[3]) {
  int off[2];
  const int idx[2] = {i0, i1};
  for (int d = 0; d < 2; d++) {
    off[d] = face[d];
    if (idx[d] < 5) off[d] = 5 - idx[d];
    else if (idx[d] > g->Ng[d] - 5 - 1) off[d] = (g->Ng[d] - 5 - 1) - idx[d];
  }
  const double d0 = fd1_shifted(g, f, 0, i0, i1, i2, off[0]);
  const double d1 = fd1_shifted(g, f, 1, i0, i1, i2, off[1]);
  return geo[1] * d0 + geo[2] * d1; /* d(xx2)/dr = 0 for SinhSymTP */
}

/* Radiation (Sommerfeld) condition on the RHS at pure outer points, innermost layer
 * first, then the inner parity copy on the RHS (apply_bcs_outerradiation_and_inner.c:369-469):
 *   rhs = -c (d_r f + (f - f_inf)/r) + k/r^3,
 *   k = r_int^3 (rhs_int + c (d_r f_int + (f_int - f_inf)/r_int)), f_inf = 0. */
void orc_apply_bcs_radiation_and_inner(const orc_grid *g, const orc_bc *bc, const double *xx0,
                                       const double *xx1, const double *xx2, const double *gfs,
                                       double *rhs) {
  const long nt = NTOT(g);
  for (int gz = 0; gz < NG; gz++)
    for (int dirn = 0; dirn < 3; dirn++) {
      const int slot = dirn + 3 * gz;
      for (int k = 0; k < bc->num_outer[slot]; k++) {
        const orc_outer *o = &bc->outer[slot][k];
        const int face[3] = {o->FACEX0, o->FACEX1, o->FACEX2};
        const int j0 = o->i0 + face[0], j1 = o->i1 + face[1], j2 = o->i2 + face[2];
        double geo[3], geo_int[3];
        radial_geometry(g, xx0[o->i0], xx1[o->i1], geo);
        radial_geometry(g, xx0[j0], xx1[j1], geo_int);
        const long p = IDX3(g, o->i0, o->i1, o->i2), q = IDX3(g, j0, j1, j2);
        for (int gf = 0; gf < 2; gf++) {
          const double *f = gfs + gf * nt;
          const double c = g->cOB[gf];
          const double dr = partial_r_f(g, f, o->i0, o->i1, o->i2, face, geo);
          const double dr_int = partial_r_f(g, f, j0, j1, j2, face, geo_int);
          const double out_int = -c * (dr_int + f[q] / geo_int[0]);
          const double kk = geo_int[0] * geo_int[0] * geo_int[0] * (rhs[gf * nt + q] - out_int);
          const double rinv = 1.0 / geo[0];
          rhs[gf * nt + p] = -c * (dr + f[p] * rinv) + kk * rinv * rinv * rinv;
        }
      }
    }
  orc_apply_bcs_inner_only(g, bc, rhs);
}

/* ---------------------------------------------------------------------------------
 * The Einstein Toolkit thorn's outer BC (NRPyEllipticET/src/conformally_flat_BBH/
 * conformally_flat_BBH_apply_bcs_sommerfeld.h). Same Sommerfeld condition as above, with
 * 6th-order derivatives and the falloff term extrapolated layer by layer from the
 * neighbour one point inward:
 *   rhs_p = -c (d_r f_p + (f_p - f_inf)/r_p)
 *           + (rhs_q + c (d_r f_q + (f_q - f_inf)/r_q)) (r_q/r_p)^n,   q = p + FACE.
 * ------------------------------------------------------------------------------- */
/* 7-point one-sided (Fornberg: -49/20, 6, -15/2, 20/3, -15/4, 6/5, -1/6) and centred
 * (3/4, -3/20, 1/60) first-derivative weights, conformally_flat_BBH_apply_bcs_sommerfeld.h:33-46 */
static const double SOM_UP[7] = {-49.0 / 20.0, 6.0, -15.0 / 2.0, 20.0 / 3.0, -15.0 / 4.0, 6.0 / 5.0, -1.0 / 6.0};
static const double SOM_C1 = 3.0 / 4.0, SOM_C2 = 3.0 / 20.0, SOM_C3 = 1.0 / 60.0;

/* conformally_flat_BBH_apply_bcs_sommerfeld.h:49-105: one-sided along the face normal, and
 * wherever idx + NGHOSTS >= Ng (backward) or idx - NGHOSTS <= 0 (forward); centred otherwise.
 * The one-sided sum runs from the point itself outward, the centred one from +3 down to -3. */
double orc_sommerfeld_fd1(const orc_grid *g, const double *f, int dir, int i0, int i1, int i2, int face) {
  const int idx = dir == 0 ? i0 : i1;
  const long st = dir == 0 ? 1 : g->Ng[0];
  const long p = IDX3(g, i0, i1, i2);
  if (abs(face) == 1 || idx + NG >= g->Ng[dir] || idx - NG <= 0) {
    int sh = face;
    if (idx + NG >= g->Ng[dir]) sh = -1;
    if (idx - NG <= 0) sh = +1;
    double s = SOM_UP[0] * f[p];
    for (int k = 1; k < 7; k++) s = fma(SOM_UP[k], f[p + k * sh * st], s);
    return (sh * s) * g->invdx[dir];
  }
  double s = SOM_C3 * f[p + 3 * st];
  s = fma(-SOM_C2, f[p + 2 * st], s);
  s = fma(SOM_C1, f[p + st], s);
  s = fma(-SOM_C1, f[p - st], s);
  s = fma(SOM_C2, f[p - 2 * st], s);
  s = fma(-SOM_C3, f[p - 3 * st], s);
  return s * g->invdx[dir];
}

/* d_r f = dxx0/dr d_0 f + dxx1/dr d_1 f (conformally_flat_BBH_contraction_term, :16-137; r does
 * not depend on xx2). The thorn's closed form for the two Jacobian factors is the inverse of
 * d(r, cos theta)/d(xx0, xx1), the same quantity radial_geometry restates from the map. */
static double sommerfeld_dr(const orc_grid *g, const double *f, int i0, int i1, int i2, const int face[3],
                            const double geo[3]) {
  const double d0 = orc_sommerfeld_fd1(g, f, 0, i0, i1, i2, face[0]);
  const double d1 = orc_sommerfeld_fd1(g, f, 1, i0, i1, i2, face[1]);
  return geo[1] * d0 + geo[2] * d1;
}

/* conformally_flat_BBH_apply_bcs_sommerfeld.h:139-222. The thorn copies inner ghost points
 * layer by layer after that layer's outer points; every inner source is an interior point or
 * a pure outer point of the same or an inner layer (the SinhSymTP maps keep i0 or move it into
 * the interior), so one copy after all layers gives the same values. */
void orc_apply_bcs_sommerfeld(const orc_grid *g, const orc_bc *bc, const double *xx0, const double *xx1,
                              const double *xx2, const double speed[2], const double finf[2],
                              const double falloff[2], const double *gfs, double *rhs) {
  (void)xx2;
  const long nt = NTOT(g);
  for (int gz = 0; gz < NG; gz++)
    for (int dirn = 0; dirn < 3; dirn++) {
      const int slot = dirn + 3 * gz;
      for (int k = 0; k < bc->num_outer[slot]; k++) {
        const orc_outer *o = &bc->outer[slot][k];
        const int face[3] = {o->FACEX0, o->FACEX1, o->FACEX2};
        const int j0 = o->i0 + face[0], j1 = o->i1 + face[1], j2 = o->i2 + face[2];
        double geo[3], geo_q[3];
        radial_geometry(g, xx0[o->i0], xx1[o->i1], geo);
        radial_geometry(g, xx0[j0], xx1[j1], geo_q);
        const long p = IDX3(g, o->i0, o->i1, o->i2), q = IDX3(g, j0, j1, j2);
        for (int gf = 0; gf < 2; gf++) {
          const double *f = gfs + gf * nt;
          double *rh = rhs + gf * nt;
          const double c = speed[gf];
          const double invr = 1.0 / geo[0];
          double val = -c * (sommerfeld_dr(g, f, o->i0, o->i1, o->i2, face, geo) + invr * (f[p] - finf[gf]));
          if (falloff[gf] > 0) {
            const double invr_q = 1.0 / geo_q[0];
            const double extrap = c * (sommerfeld_dr(g, f, j0, j1, j2, face, geo_q) + invr_q * (f[q] - finf[gf]));
            val += (rh[q] + extrap) * pow(geo_q[0] * invr, falloff[gf]);
          }
          rh[p] = val;
        }
      }
    }
  orc_apply_bcs_inner_only(g, bc, rhs);
}

static void outer_bc(const orc_grid *g, const orc_bc *bc, const double *xx0, const double *xx1,
                     const double *xx2, const double *gfs, double *rhs) {
  if (g->outer_bc == ORC_BC_SOMMERFELD6) {
    const double finf[2] = {0.0, 0.0}, falloff[2] = {3.0, 3.0};  /* :10-11 */
    orc_apply_bcs_sommerfeld(g, bc, xx0, xx1, xx2, g->cOB, finf, falloff, gfs, rhs);
  } else {
    orc_apply_bcs_radiation_and_inner(g, bc, xx0, xx1, xx2, gfs, rhs);
  }
}

/* ---------------------------------------------------------------------------------
 * RK4 method of lines, k_odd/k_even scheme (MoL_step_forward_in_time.c:7-178; Butcher
 * table RK_Butcher_Table_Dictionary.py:86-91). Stage updates cover every point,
 * ghosts included, with the reference FMA placement (:44-45, :88-89, :129-130, :170).
 * With g->outer_bc == ORC_BC_SOMMERFELD6 the thorn's BC takes the radiation BC's place, as in
 * conformally_flat_BBH_MoL_step_forward_in_time.c:27,53,80,107.
 * ------------------------------------------------------------------------------- */
void orc_mol_step(const orc_grid *g, const orc_rfm *rfm, const orc_bc *bc, const double *xx0,
                  const double *xx1, const double *xx2, const double *aux, double *y_n, double *rt,
                  double *k_odd, double *k_even, double dt) {
  const long n = 2 * NTOT(g);
  /* k1 */
  orc_rhs_eval(g, rfm, aux, y_n, k_odd);
  outer_bc(g, bc, xx0, xx1, xx2, y_n, k_odd);
  for (long i = 0; i < n; i++) {
    const double k = k_odd[i];
    rt[i] = (1.0 / 6.0) * (dt * k);
    k_odd[i] = fma(1.0 / 2.0, dt * k, y_n[i]);
  }
  /* k2 */
  orc_rhs_eval(g, rfm, aux, k_odd, k_even);
  outer_bc(g, bc, xx0, xx1, xx2, k_odd, k_even);
  for (long i = 0; i < n; i++) {
    const double k = k_even[i];
    rt[i] = fma(1.0 / 3.0, dt * k, rt[i]);
    k_even[i] = fma(1.0 / 2.0, dt * k, y_n[i]);
  }
  /* k3 */
  orc_rhs_eval(g, rfm, aux, k_even, k_odd);
  outer_bc(g, bc, xx0, xx1, xx2, k_even, k_odd);
  for (long i = 0; i < n; i++) {
    const double k = k_odd[i];
    rt[i] = fma(1.0 / 3.0, dt * k, rt[i]);
    k_odd[i] = fma(dt, k, y_n[i]);
  }
  /* k4 */
  orc_rhs_eval(g, rfm, aux, k_odd, k_even);
  outer_bc(g, bc, xx0, xx1, xx2, k_odd, k_even);
  for (long i = 0; i < n; i++) y_n[i] = y_n[i] + fma(1.0 / 6.0, dt * k_even[i], rt[i]);
}

/* L2_norm_of_gf.c:6-43: volume-weighted RMS over interior points with r < radius,
 * returning log10(1e-16 + sqrt(sum gf^2 dV / sum dV)); dV = sqrt(det gamma_hat) dxx0 dxx1 dxx2,
 * sqrt(det gamma_hat) = f0 f0' f2^2 |sin x1| / f3 for SinhSymTP. Serial sum in loop order. */
double orc_l2_norm(const orc_grid *g, double radius, const double *xx0, const double *xx1,
                   const double *xx2, const double *gf) {
  const double w = g->SINHWAA, A = g->AMAX, b = g->bScale, D = sinh_den(g);
  double S = 0.0, V = 0.0;
  for (int i2 = NG; i2 < NG + g->N[2]; i2++)
    for (int i1 = NG; i1 < NG + g->N[1]; i1++)
      for (int i0 = NG; i0 < NG + g->N[0]; i0++) {
        const double ep = exp(xx0[i0] / w), em = exp(-xx0[i0] / w);
        const double f0 = A * (ep - em) / D, f0p = A * (ep + em) / (w * D);
        const double s = sin(xx1[i1]), c = cos(xx1[i1]);
        const double f3sq = f0 * f0 + b * b, f2sq = f0 * f0 + b * b * s * s;
        const double r = sqrt(s * s * f0 * f0 + f3sq * c * c);
        const double sqrtdet = f0 * f0p * f2sq * fabs(s) / sqrt(f3sq);
        if (r < radius) {
          const double val = gf[IDX3(g, i0, i1, i2)];
          const double dV = sqrtdet * g->dxx[0] * g->dxx[1] * g->dxx[2];
          S += val * val * dV;
          V += dV;
        }
      }
  return log10(1e-16 + sqrt(S / V));
}

/* =====================================================================================
 * Einstein Toolkit consumer path: the converged u on the SinhSymTP grid -> ADM fields on
 * the evolution code's Cartesian points (NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase.c:5-200).
 * ===================================================================================== */

/* conformally_flat_BBH_get_xx_from_xyz.c:6-20: the closed-form inverse of the SinhSymTP map.
 * With R2 = x^2+y^2+z^2 and Q = sqrt((R2 - b^2)^2 + 4 b^2 (x^2+y^2)) (expanded as the
 * reference expands it):
 *   xx0 = w asinh( sqrt((R2 - b^2 + Q)/2) sinh(1/w) / AMAX )   (the SymTP radial coordinate,
 *         pulled back through the sinh stretch),
 *   xx1 = acos( 0.99999999 sign(z) sqrt((1 + (R2 - Q)/b^2)/2) ) (the factor keeps xx1 off the
 *         axis), xx2 = atan2(y, x). */
void orc_get_xx_from_xyz(double AMAX, double SINHWAA, double bScale, const double X[3], double xx[3]) {
  const double b2 = bScale * bScale;
  const double R2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  const double Q = sqrt(-4 * (X[2] * X[2]) * b2 + bScale * bScale * bScale * bScale + 2 * b2 * R2 + R2 * R2);
  const double ib2 = 1.0 / b2;
  const double sgn = (double)((X[2] > 0) - (X[2] < 0));
  xx[0] = SINHWAA * asinh(M_SQRT1_2 * sqrt(-b2 + R2 + Q) * sinh(1.0 / SINHWAA) / AMAX);
  xx[1] = acos(0.99999999 * M_SQRT1_2 * sqrt(R2 * ib2 - Q * ib2 + 1) * sgn);
  xx[2] = atan2(X[1], X[0]);
}

/* Lagrange interpolation with the molecule placement of the Cactus thorn AEILocalInterp
 * (Einstein Toolkit; "Lagrange polynomial interpolation", which
 * NRPyEllipticET_conformally_flat_BBH_interpolate_solution_to_ADMBase.c:14-55 requests with
 * "order=<interpolation_order>", param.ccl:67-75). Per axis: fp = (x - origin)/delta in grid
 * units; a molecule of order+1 points spans m = m_min..m_min+order, m_min = -floor(order/2),
 * around the nearest point (odd point count) or the point below x (even count); it is moved
 * inward where it would leave the array; points more than 1e-10 grid spacings outside the
 * array are an error. Weights are the Lagrange basis polynomials at fp - centre, each the product
 * of (t - k) times the reciprocal of its (integer) denominator. */
static int lagrange_axis(int n, double origin, double delta, int order, double x, int *first, double w[8]) {
  const double fp = (x - origin) / delta;
  if (!(fp >= -1e-10 && fp <= (double)(n - 1) + 1e-10)) return -1;
  const int mmin = -(order / 2), mmax = mmin + order;
  int c = (order % 2 == 0) ? (int)floor(fp + 0.5) : (int)floor(fp);
  if (c + mmin < 0) c = -mmin;
  if (c + mmax > n - 1) c = n - 1 - mmax;
  const double t = fp - (double)c;
  for (int m = mmin; m <= mmax; m++) {
    double v = 1.0, den = 1.0;
    for (int k = mmin; k <= mmax; k++)
      if (k != m) {
        v *= t - (double)k;
        den *= (double)(m - k);
      }
    w[m - mmin] = v * (1.0 / den);
  }
  *first = c + mmin;
  return 0;
}

int orc_lagrange_interp3(const int dims[3], const double origin[3], const double delta[3], const double *data,
                         int order, const double pt[3], double *out) {
  if (order < 1 || order > 6) return -1;
  int f[3];
  double w[3][8];
  for (int d = 0; d < 3; d++)
    if (lagrange_axis(dims[d], origin[d], delta[d], order, pt[d], &f[d], w[d])) return -1;
  double acc = 0.0;
  for (int k = 0; k <= order; k++) {
    double a1 = 0.0;
    for (int j = 0; j <= order; j++) {
      double a0 = 0.0;
      const double *row = data + (long)f[0] + (long)dims[0] * ((long)(f[1] + j) + (long)dims[1] * (long)(f[2] + k));
      for (int i = 0; i <= order; i++) a0 += w[0][i] * row[i];
      a1 += w[1][j] * a0;
    }
    acc += w[2][k] * a1;
  }
  *out = acc;
  return 0;
}

/* conformally_flat_BBH_compute_ADM_Cartesian_quantities_from_uu.c:7-103 and the two lapse
 * choices (conformally_flat_BBH_lapse_power_of_psi_initial_data.c:6-35,
 * conformally_flat_BBH_lapse_averaged_initial_data.c:6-41), at one point of the ETK grid:
 *   psi = 1 + uu + sum_n m_n/(2 r_n),  gamma_ij = psi^4 delta_ij,  beta^i = 0,
 *   K_ij = psi^-2 A_ij with the Bowen-York A_ij of orc_auxevol,
 *   alpha = psi^n (psi^n lapse) or 1/2 + (1 - F)/(2 (1 + F)), F = sum_n m_n/(2 r_n) (averaged). */
void orc_adm_point(const orc_grid *etk, int lapse, double lapse_n, double x, double y, double z, double uu,
                   double out[ORC_ADM_NOUT]) {
  const double X[3] = {x, y, z};
  double A[3][3] = {{0}}, h[2];
  for (int n = 0; n < 2; n++) {
    const double r[3] = {X[0] - etk->pos[n][0], X[1] - etk->pos[n][1], X[2] - etk->pos[n][2]};
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], rr = sqrt(r2);
    const double ir3 = 1.0 / (r2 * rr), ir5 = ir3 / r2;
    const double *P = etk->P[n], *S = etk->S[n];
    const double rP = r[0] * P[0] + r[1] * P[1] + r[2] * P[2];
    const double rS[3] = {r[1] * S[2] - r[2] * S[1], r[2] * S[0] - r[0] * S[2], r[0] * S[1] - r[1] * S[0]};
    h[n] = etk->mass[n] / (2.0 * rr);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        A[i][j] += 1.5 * (P[i] * r[j] + P[j] * r[i]) * ir3 - (i == j ? 1.5 * rP * ir3 : 0.0) +
                   1.5 * rP * r[i] * r[j] * ir5 - 3.0 * (rS[i] * r[j] + rS[j] * r[i]) * ir5;
  }
  /* sums in the reference's order: psi = ((uu + 1) + h0) + h1, psi^4 = ((psi psi) psi) psi */
  const double F = h[0] + h[1];
  const double psi = ((uu + 1.0) + h[0]) + h[1], psi4 = psi * psi * psi * psi, ipsi2 = 1.0 / (psi * psi);
  out[0] = uu;
  out[1] = lapse == ORC_LAPSE_AVERAGED ? 0.5 * (1.0 - F) / (F + 1.0) + 0.5 : pow((F + uu) + 1.0, lapse_n);
  out[2] = out[3] = out[4] = 0.0;
  static const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int c = 0; c < 6; c++) {
    out[5 + c] = ij[c][0] == ij[c][1] ? psi4 : 0.0;
    out[11 + c] = ipsi2 * A[ij[c][0]][ij[c][1]];
  }
}

/* NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase.c:35-200. The solve ran in the NRPy
 * frame; with orbital plane "xy" the relaxation driver mapped the parameter file's punctures
 * (x,y,z)_NRPy = (y,z,x)_ETK (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:84-105),
 * so the ETK-frame punctures the ADM formulas use are the inverse permutation of g's, plus
 * position_shift. Each ETK point, less the shift, is permuted the same way into the NRPy frame,
 * mapped to (xx0,xx1,xx2) (retried once, moved by nrpy_epsilon on every axis, if that gives a
 * NaN) and u is interpolated there. */
int orc_initialize_admbase(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2,
                           const double *u, const orc_adm_opts *o, long n, const double *x, const double *y,
                           const double *z, double *const out[ORC_ADM_NOUT]) {
  orc_grid etk = *g;
  const int xy = o->orbital_plane == ORC_PLANE_XY;
  for (int p = 0; p < 2; p++) {
    double *v[3] = {etk.pos[p], etk.P[p], etk.S[p]};
    const double *w[3] = {g->pos[p], g->P[p], g->S[p]};
    for (int q = 0; q < 3; q++) {
      if (xy) {  /* ETK (x,y,z) = NRPy (z,x,y) */
        v[q][0] = w[q][2];
        v[q][1] = w[q][0];
        v[q][2] = w[q][1];
      }
    }
    for (int d = 0; d < 3; d++) etk.pos[p][d] += o->position_shift[d];
  }
  const int dims[3] = {g->Ng[0], g->Ng[1], g->Ng[2]};
  const double origin[3] = {xx0[0], xx1[0], xx2[0]};
  const double delta[3] = {xx0[1] - xx0[0], xx1[1] - xx1[0], xx2[1] - xx2[0]};
  int rc = 0;
  for (long i = 0; i < n; i++) {
    const double s[3] = {x[i] - o->position_shift[0], y[i] - o->position_shift[1], z[i] - o->position_shift[2]};
    double X[3], xs[3];
    if (xy) { X[0] = s[1]; X[1] = s[2]; X[2] = s[0]; }
    else { X[0] = s[0]; X[1] = s[1]; X[2] = s[2]; }
    orc_get_xx_from_xyz(g->AMAX, g->SINHWAA, g->bScale, X, xs);
    if (isnan(xs[0] * xs[1] * xs[2])) {
      for (int d = 0; d < 3; d++) X[d] += o->nrpy_epsilon;
      orc_get_xx_from_xyz(g->AMAX, g->SINHWAA, g->bScale, X, xs);
      if (isnan(xs[0] * xs[1] * xs[2])) return -1;
    }
    double uu;
    if (orc_lagrange_interp3(dims, origin, delta, u, o->interp_order, xs, &uu)) return -2;
    double v[ORC_ADM_NOUT];
    orc_adm_point(&etk, o->initial_lapse, o->lapse_exponent, x[i], y[i], z[i], uu, v);
    for (int c = 0; c < ORC_ADM_NOUT; c++) out[c][i] = v[c];
  }
  return rc;
}

/* conformally_flat_BBH_set_optimum_eta_damping.c:3-27: the relaxation time-scale ratio of
 * arXiv:2111.02424 eq. (B6),
 *   TRC = dt/(C_0 dxx1) atanh(A/sqrt(A^2 + b^2 cos^2(dxx1/2))) cos(dxx1/2) cos(dxx2/2),
 * and the fitted optimum damping of its eq. (36), eta = 6.9 TRC^-0.79 - 0.7. */
double orc_compute_TRC(double A, double b, double C_0, double dt, double dxx1, double dxx2) {
  const double c1 = cos(dxx1 / 2.0), c2 = cos(dxx2 / 2.0);
  const double t = A / sqrt(A * A + b * b * c1 * c1);
  return dt / (C_0 * dxx1) * atanh(t) * c1 * c2;
}
double orc_optimum_eta(double AMAX, double bScale, double dxx1, double dxx2, double dt, double CFL) {
  return 6.9 * pow(orc_compute_TRC(AMAX, bScale, CFL, dt, dxx1, dxx2), -0.79) - 0.7;
}
