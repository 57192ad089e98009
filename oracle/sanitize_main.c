/*
 * sanitize_main.c -- drives every function of the C restatement (oracle.c) once on a small
 * SinhSymTP grid, for a build under AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`, run by tests/test_oracle_sanitize.py). TEST INFRASTRUCTURE.
 *
 * Exits 0 when every call returned normally and every result is finite; a sanitizer report
 * aborts the process with a nonzero status.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double *dz(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

int main(void) {
  orc_grid g;
  memset(&g, 0, sizeof(g));
  g.N[0] = 16; g.N[1] = 12; g.N[2] = 6;
  g.AMAX = 1e6; g.SINHWAA = 0.07; g.bScale = 5.0; g.eta = 20.0;
  g.mass[0] = 0.52; g.mass[1] = 0.39;
  g.pos[0][2] = 5.0; g.pos[1][2] = -5.0;
  g.P[0][0] = 0.095; g.P[1][0] = -0.095;
  g.S[0][1] = 0.095; g.S[1][1] = -0.09;
  const int Ng[3] = {g.N[0] + 2 * ORC_NGHOSTS, g.N[1] + 2 * ORC_NGHOSTS, g.N[2] + 2 * ORC_NGHOSTS};
  const long ntot = (long)Ng[0] * Ng[1] * Ng[2];
  double *xe[3], *xx[3];
  for (int d = 0; d < 3; d++) { xe[d] = dz(Ng[d]); xx[d] = dz(Ng[d]); }

  /* eigen grid -> bcstruct, then the SinhSymTP grid */
  if (orc_set_grid(&g, 1, xe[0], xe[1], xe[2])) return 10;
  orc_bc bc;
  memset(&bc, 0, sizeof(bc));
  bc.inner = (orc_inner *)calloc((size_t)ntot, sizeof(orc_inner));
  for (int s = 0; s < ORC_NGHOSTS * 3; s++) bc.outer[s] = (orc_outer *)calloc((size_t)ntot, sizeof(orc_outer));
  if (orc_bcstruct(&g, xe[0], xe[1], xe[2], &bc)) return 11;
  if (orc_set_grid(&g, 0, xx[0], xx[1], xx[2])) return 12;
  const double dt = orc_find_timestep(&g, xx[0], xx[1], xx[2]);

  orc_rfm rfm;
  double **f = (double **)&rfm;
  for (int k = 0; k < 15; k++) {
    const long n = k < 4 ? Ng[0] : k < 7 ? Ng[1] : k < 12 ? (long)Ng[0] * Ng[1] : Ng[0];
    f[k] = dz(n);
  }
  orc_rfm_define(&g, xx[0], xx[1], &rfm);
  double *aux = dz(3 * ntot);
  orc_auxevol(&g, xx[0], xx[1], xx[2], aux);
  orc_wavespeed(&g, 0.7, dt, xx[0], xx[1], xx[2], aux);
  g.cOB[0] = g.cOB[1] = orc_wavespeed_at_OB(&g, aux);

  /* RHS, BCs, three RK4 steps, residual and L2 */
  double *yn = dz(2 * ntot), *rt = dz(2 * ntot), *ko = dz(2 * ntot), *ke = dz(2 * ntot), *res = dz(ntot);
  for (long i = 0; i < 2 * ntot; i++) yn[i] = 1e-3 * sin(0.001 * (double)i);
  orc_apply_bcs_inner_only(&g, &bc, yn);
  orc_rhs_eval(&g, &rfm, aux, yn, ko);
  orc_apply_bcs_radiation_and_inner(&g, &bc, xx[0], xx[1], xx[2], yn, ko);
  for (int s = 0; s < 3; s++) orc_mol_step(&g, &rfm, &bc, xx[0], xx[1], xx[2], aux, yn, rt, ko, ke, dt);
  /* the thorn's Sommerfeld BC, alone and inside the RK4 step */
  {
    const double sp[2] = {4.4793122676309114e+06, 4.4793122676309114e+06}, fi[2] = {0, 0}, fo[2] = {3, 3};
    orc_apply_bcs_sommerfeld(&g, &bc, xx[0], xx[1], xx[2], sp, fi, fo, yn, ko);
    g.outer_bc = ORC_BC_SOMMERFELD6;
    orc_mol_step(&g, &rfm, &bc, xx[0], xx[1], xx[2], aux, yn, rt, ko, ke, dt);
    g.outer_bc = ORC_BC_RADIATION;
  }
  orc_residual(&g, &rfm, aux, yn, res);
  const double l2 = orc_l2_norm(&g, 100.0, xx[0], xx[1], xx[2], res);
  double w = 0;
  for (int o = -5; o <= 5; o++)
    for (int j = 0; j < 11; j++) w += orc_fd1_weight(o, j);

  /* consumer path: Cartesian points -> coordinates -> interpolation -> ADM fields */
  const long np = 500;
  double *px = dz(np), *py = dz(np), *pz = dz(np), *out[ORC_ADM_NOUT];
  for (int c = 0; c < ORC_ADM_NOUT; c++) out[c] = dz(np);
  for (long i = 0; i < np; i++) {
    px[i] = 10.0 * sin(1.3 * (double)i);
    py[i] = 10.0 * cos(0.7 * (double)i);
    pz[i] = 7.0 * sin(0.31 * (double)i + 0.2);
  }
  orc_adm_opts o = {ORC_PLANE_XY, {0.1, -0.2, 0.3}, ORC_LAPSE_AVERAGED, -2.0, 1e-8, 4};
  const int rc = orc_initialize_admbase(&g, xx[0], xx[1], xx[2], yn, &o, np, px, py, pz, out);
  const double eta = orc_optimum_eta(g.AMAX, g.bScale, g.dxx[1], g.dxx[2], dt, 0.7);

  double sum = 0;
  for (long i = 0; i < 2 * ntot; i++) sum += fabs(yn[i]);
  for (long i = 0; i < np; i++) sum += fabs(out[0][i]) + fabs(out[1][i]) + fabs(out[16][i]);
  printf("dt %.6e l2 %.6f w %.3e adm_rc %d eta %.4f sum %.6e\n", dt, l2, w, rc, eta, sum);
  const int ok = rc == 0 && isfinite(dt) && isfinite(l2) && isfinite(sum) && isfinite(eta);

  for (int d = 0; d < 3; d++) { free(xe[d]); free(xx[d]); }
  free(bc.inner);
  for (int s = 0; s < ORC_NGHOSTS * 3; s++) free(bc.outer[s]);
  for (int k = 0; k < 15; k++) free(f[k]);
  free(aux); free(yn); free(rt); free(ko); free(ke); free(res);
  free(px); free(py); free(pz);
  for (int c = 0; c < ORC_ADM_NOUT; c++) free(out[c]);
  return ok ? 0 : 1;
}
