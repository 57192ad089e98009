/*
 * oracle.h -- CPU restatement of NRPyElliptic's hyperbolic-relaxation hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. The product
 * (nrpyelliptic_amd/libnrpyell_hip.so) never links, loads or calls it.
 *
 * Every function restates one reference function (NRPyElliptic_Ccodes/<file>:<line>)
 * in plain C.  Data layout follows the reference: gridfunction-major SoA with i0
 * fastest, IDX4S(g,i0,i1,i2) = i0 + Ng0*(i1 + Ng1*(i2 + Ng2*g)), ghosts included
 * (NRPy_basic_defines.h:247-250).
 *
 * Parity status of each function (how it is pinned) is listed in oracle.c's header.
 */
#ifndef NRPYELL_ORACLE_H
#define NRPYELL_ORACLE_H

#include <stdint.h>

#define ORC_NGHOSTS 5

/* Grid + physics parameters (the subset of the reference paramstruct the path reads). */
typedef struct {
  int N[3];          /* interior points per direction (Nxx0..2)                  */
  int Ng[3];         /* N + 2*NGHOSTS                                            */
  double dxx[3];     /* grid spacing of the grid last set by orc_set_grid         */
  double invdx[3];
  double AMAX, SINHWAA, bScale;
  double eta;        /* eta_damping                                              */
  double mass[2];    /* bare masses                                              */
  double pos[2][3];  /* puncture positions                                       */
  double P[2][3];    /* linear momenta                                           */
  double S[2][3];    /* spins                                                    */
  double cOB[2];     /* wavespeed used by the radiation BC for u and v           */
  int outer_bc;      /* ORC_BC_RADIATION (default) or ORC_BC_SOMMERFELD6          */
} orc_grid;

/* Outer boundary condition of orc_mol_step. */
#define ORC_BC_RADIATION 0   /* NRPyElliptic: 11-point 10th-order stencils, k/r^3 extrapolation */
#define ORC_BC_SOMMERFELD6 1 /* Einstein Toolkit thorn: 7-point 6th-order stencils, radial
                                falloff power 3, wave speed cOB (the thorn hard-codes it)      */

/* Same member order as the reference rfm_struct (NRPy_basic_defines.h:135-151). */
typedef struct {
  double *f0, *f0_D0, *f0_DD00, *f0_DDD000;
  double *f1, *f1_D1, *f1_DD11;
  double *f2, *f2_D0, *f2_D1, *f2_DD00, *f2_DD11;
  double *f3, *f3_D0, *f3_DD00;
} orc_rfm;

/* Ghost-zone tables in the reference's own record layout (NRPy_basic_defines.h:164-178). */
typedef struct { int dstpt, srcpt; int8_t parity[10]; } orc_inner;
typedef struct { short i0, i1, i2; int8_t FACEX0, FACEX1, FACEX2; } orc_outer;

/* Outer lists are stored per (which_gz, dirn) slot dirn + 3*which_gz, the slot the
 * reference reads them from (apply_bcs_outerradiation_and_inner.c:441-447). */
typedef struct {
  int num_inner;
  orc_inner *inner;                  /* caller-allocated, >= number of ghost points */
  int num_outer[ORC_NGHOSTS * 3];
  orc_outer *outer[ORC_NGHOSTS * 3]; /* each caller-allocated, >= face size          */
} orc_bc;

/* AUXEVOL gridfunction ids, NRPy_basic_defines.h:230-233 */
#define ORC_ADD_TIMES_AUU 0
#define ORC_PSI_BACKGROUND 1
#define ORC_WAVESPEED 2

int    orc_set_grid(orc_grid *g, int eigen, double *xx0, double *xx1, double *xx2);
void   orc_rfm_define(const orc_grid *g, const double *xx0, const double *xx1, orc_rfm *rfm);
void   orc_auxevol(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2, double *aux);
double orc_find_timestep(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2);
void   orc_wavespeed(const orc_grid *g, double CFL, double dt, const double *xx0, const double *xx1,
                     const double *xx2, double *aux);
double orc_wavespeed_at_OB(const orc_grid *g, const double *aux);
int    orc_bcstruct(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2, orc_bc *bc);
void   orc_rhs_eval(const orc_grid *g, const orc_rfm *rfm, const double *aux, const double *in, double *rhs);
void   orc_residual(const orc_grid *g, const orc_rfm *rfm, const double *aux, const double *in, double *res);
void   orc_apply_bcs_inner_only(const orc_grid *g, const orc_bc *bc, double *gfs);
void   orc_apply_bcs_radiation_and_inner(const orc_grid *g, const orc_bc *bc, const double *xx0,
                                         const double *xx1, const double *xx2, const double *gfs,
                                         double *rhs);
/* The Einstein Toolkit thorn's outer BC (conformally_flat_BBH_apply_bcs_sommerfeld.h:139-222)
 * followed by the inner parity copy on the RHS; speed, f_inf and the radial falloff power per
 * evolved gridfunction. */
void   orc_apply_bcs_sommerfeld(const orc_grid *g, const orc_bc *bc, const double *xx0, const double *xx1,
                                const double *xx2, const double speed[2], const double finf[2],
                                const double falloff[2], const double *gfs, double *rhs);
/* 7-point first derivative of the thorn along dir (0: x0, 1: x1) at (i0,i1,i2) with face
 * direction `face` (conformally_flat_BBH_apply_bcs_sommerfeld.h:16-105); exposed for tests. */
double orc_sommerfeld_fd1(const orc_grid *g, const double *f, int dir, int i0, int i1, int i2, int face);
void   orc_mol_step(const orc_grid *g, const orc_rfm *rfm, const orc_bc *bc, const double *xx0,
                    const double *xx1, const double *xx2, const double *aux, double *y_n, double *rt,
                    double *k_odd, double *k_even, double dt);
double orc_l2_norm(const orc_grid *g, double radius, const double *xx0, const double *xx1,
                   const double *xx2, const double *gf);
/* FD1 weight for 11-point first derivative with stencil offset `offset` (-5..5), tap j=0..10
 * at i0+offset-5+j; exposed for the closed-form test. */
double orc_fd1_weight(int offset, int j);

/* ---- Einstein Toolkit consumer path (NRPyEllipticET/src/conformally_flat_BBH/) ---------- */

/* Options of NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase (param.ccl:39-88,142). */
#define ORC_PLANE_XY 0      /* "xy"/"yx": (x,y,z)_NRPy = (y,z,x)_ETK */
#define ORC_PLANE_XZ 1      /* "xz"/"zx": same axes */
#define ORC_LAPSE_PSI_N 0   /* "NRPyEllipticET-psi^n"   */
#define ORC_LAPSE_AVERAGED 1 /* "NRPyEllipticET-averaged" */
#define ORC_ADM_NOUT 17     /* uu, alp, betax..z, gxx gxy gxz gyy gyz gzz, kxx kxy kxz kyy kyz kzz */
typedef struct {
  int orbital_plane;
  double position_shift[3];
  int initial_lapse;
  double lapse_exponent;
  double nrpy_epsilon;
  int interp_order;          /* Lagrange order, 1..6 */
} orc_adm_opts;

void   orc_get_xx_from_xyz(double AMAX, double SINHWAA, double bScale, const double X[3], double xx[3]);
/* 3-D tensor-product Lagrange interpolation on a uniform grid; 0 on success, -1 outside the grid. */
int    orc_lagrange_interp3(const int dims[3], const double origin[3], const double delta[3],
                            const double *data, int order, const double pt[3], double *out);
/* ADM quantities at one ETK point from uu (punctures of `etk` in the ETK frame, shifted);
 * out[1..16] = alp, beta (3), gamma_ij (6), K_ij (6); out[0] = uu. */
void   orc_adm_point(const orc_grid *etk, int lapse, double lapse_n, double x, double y, double z, double uu,
                     double out[ORC_ADM_NOUT]);
/* The whole consumer path on n points; g holds the NRPy-frame grid and punctures the solve ran
 * with. Returns 0, -1 (unrecoverable NaN coordinate) or -2 (point outside the grid). */
int    orc_initialize_admbase(const orc_grid *g, const double *xx0, const double *xx1, const double *xx2,
                              const double *u, const orc_adm_opts *o, long n, const double *x,
                              const double *y, const double *z, double *const out[ORC_ADM_NOUT]);
double orc_compute_TRC(double A, double b, double C_0, double dt, double dxx1, double dxx2);
double orc_optimum_eta(double AMAX, double bScale, double dxx1, double dxx2, double dt, double CFL);

#endif
