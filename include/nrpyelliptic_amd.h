/*
 * nrpyelliptic_amd.h -- C ABI of libnrpyell_hip.so, the MI355X (gfx950) implementation
 * of NRPyElliptic's hyperbolic-relaxation solve for binary-puncture initial data.
 *
 * Part 1 is the drop-in surface: the structs and functions of the reference driver
 * (NRPyElliptic_Ccodes/NRPy_basic_defines.h, NRPy_function_prototypes.h) with the same
 * names, member order, prototypes and argument meaning, so a caller of the reference
 * links against this library instead.  The one change of contract is WHERE arrays live:
 *   - gridfunction arrays (MoL_gridfunctions_struct members, auxevol_gfs, in_gfs,
 *     rhs_gfs, residual_gf, in_gf) and rfm_struct arrays are DEVICE pointers, allocated
 *     by this library's MoL_malloc_* / rfm_precompute_rfmstruct_malloc;
 *   - xx[3] coordinate arrays and bc_struct tables stay in HOST memory, exactly as the
 *     reference lays them out (bcstruct_set_up output is byte-identical).
 * Every entry point is synchronous on return, like the reference's OpenMP functions.
 * Fatal errors print to stderr and exit(1), as the reference does
 * (bcstruct_set_up.c:178-184,272-277); HIP errors are reported the same way.
 *
 * Part 2 is this library's own API (nrpyell_*): a resident solver with hipGraph-captured
 * RK4 steps and on-device residual checks, FP32->FP64 mixed precision, a batched
 * multi-configuration solver with an autotuner, and the learned initial guess.
 * None of it has a counterpart in the reference; it returns error codes instead of
 * exiting.
 */
#ifndef NRPYELLIPTIC_AMD_H
#define NRPYELLIPTIC_AMD_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
#define NRPYELL_RESTRICT /* restrict does not change the ABI; omitted for C++ callers */
extern "C" {
#else
#define NRPYELL_RESTRICT restrict
#endif

#define REAL double
#define NGHOSTS 5
#define NUM_EVOL_GFS 2
#define UUGF 0
#define VVGF 1
#define NUM_AUXEVOL_GFS 3
#define ADD_TIMES_AUUGF 0
#define PSI_BACKGROUNDGF 1
#define WAVESPEEDGF 2
#define EXTRAPOLATION_OUTER_BCS 0
#define RADIATION_OUTER_BCS 1

/* ---- Part 1: reference ABI (NRPy_basic_defines.h) ------------------------------------ */

/* NRPy_basic_defines.h:40-120. Member order is the ABI. */
typedef struct __paramstruct__ {
  REAL AMAX;
  REAL Cart_CoM_offsetx, Cart_CoM_offsety, Cart_CoM_offsetz;
  REAL Cart_originx, Cart_originy, Cart_originz;
  REAL P0_x, P0_y, P0_z;
  REAL P1_x, P1_y, P1_z;
  REAL S0_x, S0_y, S0_z;
  REAL S1_x, S1_y, S1_z;
  REAL SINHWAA;
  REAL bScale;
  REAL bare_mass_0, bare_mass_1;
  REAL dxx0, dxx1, dxx2;
  REAL eta_damping;
  REAL f0_of_xx0, f0_of_xx0__D0, f0_of_xx0__DD00, f0_of_xx0__DDD000;
  REAL f1_of_xx1, f1_of_xx1__D1, f1_of_xx1__DD11, f1_of_xx1__DDD111;
  REAL f2_of_xx0_xx1, f2_of_xx0_xx1__D0, f2_of_xx0_xx1__D1, f2_of_xx0_xx1__DD00, f2_of_xx0_xx1__DD11;
  REAL f2_of_xx1, f2_of_xx1__D1, f2_of_xx1__DD11, f2_of_xx1__DDD111;
  REAL f3_of_xx0, f3_of_xx0__D0, f3_of_xx0__DD00;
  REAL f4_of_xx2, f4_of_xx2__D2, f4_of_xx2__DD22;
  REAL invdx0, invdx1, invdx2;
  REAL puncture_0_x, puncture_0_y, puncture_0_z;
  REAL puncture_1_x, puncture_1_y, puncture_1_z;
  REAL time;
  REAL uu_wavespeed_at_OB, vv_wavespeed_at_OB;
  REAL xx0, xx1, xx2;
  REAL xxmax0, xxmax1, xxmax2;
  REAL xxmin0, xxmin1, xxmin2;
  int Nxx0, Nxx1, Nxx2;
  int Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1, Nxx_plus_2NGHOSTS2;
  int has_outer_boundary;
  int outer_bc_type;
} paramstruct;

/* NRPy_basic_defines.h:135-151 (device arrays here). */
typedef struct __rfmstruct__ {
  REAL *NRPYELL_RESTRICT f0_of_xx0;
  REAL *NRPYELL_RESTRICT f0_of_xx0__D0;
  REAL *NRPYELL_RESTRICT f0_of_xx0__DD00;
  REAL *NRPYELL_RESTRICT f0_of_xx0__DDD000;
  REAL *NRPYELL_RESTRICT f1_of_xx1;
  REAL *NRPYELL_RESTRICT f1_of_xx1__D1;
  REAL *NRPYELL_RESTRICT f1_of_xx1__DD11;
  REAL *NRPYELL_RESTRICT f2_of_xx0_xx1;
  REAL *NRPYELL_RESTRICT f2_of_xx0_xx1__D0;
  REAL *NRPYELL_RESTRICT f2_of_xx0_xx1__D1;
  REAL *NRPYELL_RESTRICT f2_of_xx0_xx1__DD00;
  REAL *NRPYELL_RESTRICT f2_of_xx0_xx1__DD11;
  REAL *NRPYELL_RESTRICT f3_of_xx0;
  REAL *NRPYELL_RESTRICT f3_of_xx0__D0;
  REAL *NRPYELL_RESTRICT f3_of_xx0__DD00;
} rfm_struct;

/* NRPy_basic_defines.h:164-195 (host tables). */
typedef struct __innerpt_bc_struct__ {
  int dstpt;
  int srcpt;
  int8_t parity[10];
} innerpt_bc_struct;

typedef struct __outerpt_bc_struct__ {
  short i0, i1, i2;
  int8_t FACEX0, FACEX1, FACEX2;
} outerpt_bc_struct;

typedef struct __bc_info_struct__ {
  int num_inner_boundary_points;
  int num_pure_outer_boundary_points[NGHOSTS][3];
  int bc_loop_bounds[NGHOSTS][6][6];
} bc_info_struct;

typedef struct __bc_struct__ {
  innerpt_bc_struct *NRPYELL_RESTRICT inner_bc_array;
  outerpt_bc_struct *NRPYELL_RESTRICT pure_outer_bc_array[NGHOSTS * 3];
  bc_info_struct bc_info;
} bc_struct;

/* NRPy_basic_defines.h:206-213 (device arrays here). */
typedef struct __MoL_gridfunctions_struct__ {
  REAL *NRPYELL_RESTRICT y_n_gfs;
  REAL *NRPYELL_RESTRICT y_nplus1_running_total_gfs;
  REAL *NRPYELL_RESTRICT k_odd_gfs;
  REAL *NRPYELL_RESTRICT k_even_gfs;
  REAL *NRPYELL_RESTRICT auxevol_gfs;
  REAL *NRPYELL_RESTRICT diagnostic_output_gfs;
} MoL_gridfunctions_struct;

/* NRPy_basic_defines.h:263-269 */
typedef struct __griddata__ {
  paramstruct params;
  REAL *NRPYELL_RESTRICT xx[3];
  bc_struct bcstruct;
  rfm_struct rfmstruct;
  MoL_gridfunctions_struct gridfuncs;
} griddata_struct;

/* ---- Part 1: reference functions (NRPy_function_prototypes.h) ------------------------- */

/* NRPy_function_prototypes.h:20 / set_Cparameters_to_default.c:5 */
void set_Cparameters_to_default(paramstruct *NRPYELL_RESTRICT params);
/* :11 / set_Nxx_dxx_invdx_params__and__xx.c:5. xx[] are host arrays (malloc). */
void set_Nxx_dxx_invdx_params__and__xx(const int EigenCoord, const int Nxx[3], paramstruct *NRPYELL_RESTRICT params,
                                       REAL *NRPYELL_RESTRICT xx[3]);
/* :6 / bcstruct_set_up.c:350. Host tables, byte-identical to the reference's. */
void bcstruct_set_up(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT xx[3],
                     bc_struct *NRPYELL_RESTRICT bcstruct);
/* :9 / find_timestep.c:5 */
REAL find_timestep(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT xx[3], const REAL CFL_FACTOR);
/* :1-4 / MoL_malloc_y_n_gfs.c, MoL_malloc_non_y_n_gfs.c, MoL_free_memory_*.c: device memory */
void MoL_malloc_y_n_gfs(const paramstruct *NRPYELL_RESTRICT params, MoL_gridfunctions_struct *NRPYELL_RESTRICT gridfuncs);
void MoL_free_memory_y_n_gfs(const paramstruct *NRPYELL_RESTRICT params,
                             MoL_gridfunctions_struct *NRPYELL_RESTRICT gridfuncs);
void MoL_malloc_non_y_n_gfs(const paramstruct *NRPYELL_RESTRICT params,
                            MoL_gridfunctions_struct *NRPYELL_RESTRICT gridfuncs);
void MoL_free_memory_non_y_n_gfs(const paramstruct *NRPYELL_RESTRICT params,
                                 MoL_gridfunctions_struct *NRPYELL_RESTRICT gridfuncs);
/* :17-19 / rfm_precompute_rfmstruct_{malloc,define,freemem}.c: device arrays */
void rfm_precompute_rfmstruct_malloc(const paramstruct *NRPYELL_RESTRICT params, rfm_struct *NRPYELL_RESTRICT rfmstruct);
void rfm_precompute_rfmstruct_define(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT xx[3],
                                     rfm_struct *NRPYELL_RESTRICT rfmstruct);
void rfm_precompute_rfmstruct_freemem(const paramstruct *NRPYELL_RESTRICT params, rfm_struct *NRPYELL_RESTRICT rfmstruct);
/* :24 / initial_guess_all_points.c:7: u = v = 0 at all points (device) */
void initial_guess_all_points(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT xx[3],
                              REAL *NRPYELL_RESTRICT in_gfs);
/* :27 / auxevol_gfs_all_points.c:6: psi_bg and AA^2 at all points (device) */
void auxevol_gfs_all_points(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT xx[3],
                            REAL *NRPYELL_RESTRICT auxevol_gfs);
/* :28 / wavespeed_gf_all_points.c:6 (device) */
void wavespeed_gf_all_points(const paramstruct *NRPYELL_RESTRICT params, const REAL CFL_FACTOR, const REAL dt,
                             REAL *NRPYELL_RESTRICT xx[3], REAL *NRPYELL_RESTRICT auxevol_gfs);
/* :34 / compute_wavespeed_at_OB.c:7 */
REAL compute_wavespeed_at_OB(const paramstruct *NRPYELL_RESTRICT params, REAL *NRPYELL_RESTRICT auxevol_gfs);
/* :26 / rhs_eval.c:7. HIP kernel; interior points of rhs_gfs only. */
void rhs_eval(const paramstruct *NRPYELL_RESTRICT params, const rfm_struct *NRPYELL_RESTRICT rfmstruct,
              const REAL *NRPYELL_RESTRICT auxevol_gfs, const REAL *NRPYELL_RESTRICT in_gfs,
              REAL *NRPYELL_RESTRICT rhs_gfs);
/* :7-10 / apply_bcs_outerradiation_and_inner.c:410. HIP kernels. */
void apply_bcs_outerradiation_and_inner(const paramstruct *NRPYELL_RESTRICT params,
                                        const bc_struct *NRPYELL_RESTRICT bcstruct, REAL *NRPYELL_RESTRICT xx[3],
                                        const REAL custom_wavespeed[NUM_EVOL_GFS],
                                        const REAL custom_f_infinity[NUM_EVOL_GFS], REAL *NRPYELL_RESTRICT gfs,
                                        REAL *NRPYELL_RESTRICT rhs_gfs);
/* :8 / apply_bcs_inner_only.c:11. HIP gather kernel. */
void apply_bcs_inner_only(const paramstruct *NRPYELL_RESTRICT params, const bc_struct *NRPYELL_RESTRICT bcstruct,
                          REAL *NRPYELL_RESTRICT gfs);
/* :5 / MoL_step_forward_in_time.c:7. One RK4 step of fused HIP kernels (two launches of two
 * substeps where the grid allows, else four), replayed as one captured graph per grid.
 * Precondition (always true for data produced by this path or the reference's zero
 * initial guess): y_n ghost values equal their parity images. After the call y_n is
 * exact at every point; k_odd/k_even/y_nplus1_running_total are scratch. */
void MoL_step_forward_in_time(griddata_struct *NRPYELL_RESTRICT griddata, const REAL dt);
/* :29 / residual_all_points.c:7. HIP kernel; interior points of residual_gf only. */
void residual_all_points(const paramstruct *NRPYELL_RESTRICT params, const rfm_struct *NRPYELL_RESTRICT rfmstruct,
                         const REAL *NRPYELL_RESTRICT auxevol_gfs, const REAL *NRPYELL_RESTRICT in_gfs,
                         REAL *NRPYELL_RESTRICT residual_gf);
/* :30 / L2_norm_of_gf.c:6. Deterministic two-pass HIP reduction; returns log10(1e-16+L2). */
REAL L2_norm_of_gf(const paramstruct *NRPYELL_RESTRICT params, const int gf_index, const REAL integration_radius,
                   REAL *NRPYELL_RESTRICT xx[3], const REAL *NRPYELL_RESTRICT in_gf);
/* :31-32 / gridfunction_z_axis.c:6, gridfunction_xz_plane.c:6: D2H copy, same text format. */
void gridfunction_z_axis(const paramstruct *NRPYELL_RESTRICT params, const int gf_index, REAL *NRPYELL_RESTRICT xx[3],
                         const REAL *NRPYELL_RESTRICT in_gf, FILE *NRPYELL_RESTRICT outfile);
void gridfunction_xz_plane(const paramstruct *NRPYELL_RESTRICT params, const int gf_index, REAL *NRPYELL_RESTRICT xx[3],
                           const REAL *NRPYELL_RESTRICT in_gf, FILE *NRPYELL_RESTRICT outfile);
/* :33 / print_puncture_parameters.c:6 */
void print_puncture_parameters(const paramstruct *NRPYELL_RESTRICT params);

/* ---- Part 2: this library's API (no reference counterpart) ----------------------------- */

#define NRPYELL_OK 0
#define NRPYELL_EINVAL -1
#define NRPYELL_EHIP -2
#define NRPYELL_EGHOSTMAP -3
#define NRPYELL_ENOTCONVERGED -4
#define NRPYELL_ENAN -5

/* L2_norm_of_gf that allocates nothing. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code around HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and result as L2_norm_of_gf (the log10 L2
 * goes to *log10_l2); the volume weights are computed on the device each call (device
 * exp/sin/cos: agrees with L2_norm_of_gf to rounding, not bit for bit). Every device buffer it
 * uses lives in one fixed array of NRPYELL_NOALLOC_SCRATCH_BYTES per GPU, held by the library
 * until the process ends; a grid that needs more
 * ((G0 + G1 + G0*G1 + ceil(Nxx0/64)*ceil(Nxx1/4)*Nxx2 + 1) doubles, G = Nxx + 10) is refused with
 * NRPYELL_EINVAL. Calls that overlap on one GPU share that array and race: serialise them.
 * Synchronous; NRPYELL_OK, NRPYELL_EINVAL or NRPYELL_EHIP. */
#define NRPYELL_NOALLOC_SCRATCH_BYTES (2u << 20)
int nrpyell_L2_norm_of_gf_noalloc(const paramstruct *params, const int gf_index, const REAL integration_radius,
                                  REAL *xx[3], const REAL *in_gf, REAL *log10_l2);

/* Library/device plumbing used by the Python host layer and tests. */
int nrpyell_device_synchronize(void);
int nrpyell_malloc_device(void **ptr, size_t bytes);
int nrpyell_free_device(void *ptr);
int nrpyell_memcpy_h2d(void *dst, const void *src, size_t bytes);
int nrpyell_memcpy_d2h(void *dst, const void *src, size_t bytes);
int nrpyell_memset_device(void *ptr, int value, size_t bytes);
const char *nrpyell_last_error(void);
/* Same tables as bcstruct_set_up but returns NRPYELL_EGHOSTMAP instead of exiting. */
int nrpyell_bcstruct_set_up_checked(const paramstruct *params, REAL *xx[3], bc_struct *bcstruct);
void nrpyell_bcstruct_free(bc_struct *bcstruct);
/* Checks that the closed-form SinhSymTP reflection map the fused kernels use to write ghost
 * zones reproduces this bcstruct table exactly (host only). NRPYELL_OK or NRPYELL_EGHOSTMAP. */
int nrpyell_ghostmap_check(const paramstruct *params, const bc_struct *bcstruct);

/* Resident single-configuration solver. */
typedef struct nrpyell_solver nrpyell_solver;

typedef struct {
  int precision;          /* 64: FP64 throughout; 32: FP32 bulk then FP64 finish          */
  int check_every;        /* residual + L2 every this many RK4 steps (>=1)                  */
  int steps_per_graph;    /* RK4 steps captured in one hipGraph launch (>=1)                */
  double log10_target;    /* stop when log10 L2 <= this (e.g. -10)                           */
  int max_steps;          /* give up after this many steps (NRPYELL_ENOTCONVERGED)          */
  double fp32_switch_log10;   /* precision 32: switch to FP64 at/below this log10 L2         */
  int fp32_plateau_steps;     /* precision 32: or after this many FP32 steps without a new
                                 lowest residual (the FP32 state has reached its floor)      */
  double integration_radius;  /* L2 mask radius, 100 in the reference driver (main.c:150) */
  int tile_i0;            /* i0 width of the fused substep kernel's 256-thread tile: 16, 32
                             or 64 (0 = default); the i1 height is 256 / tile_i0            */
  int planes;             /* i2 planes per thread: 1, or 2 with tile_i0 32/64 (0 = 2 where
                             the tile allows it, else 1)                                    */
  int outer_bc;           /* NRPYELL_OUTER_BC_RADIATION: apply_bcs_outerradiation_and_inner;
                             NRPYELL_OUTER_BC_SOMMERFELD6: the Einstein Toolkit thorn's BC
                             (nrpyell_apply_bcs_sommerfeld with f_infinity 0, falloff 3)      */
  double outer_bc_speed;  /* wave speed of the outer BC; 0 = the BC's default: the grid's
                             uu/vv_wavespeed_at_OB (radiation), NRPYELL_ET_SOMMERFELD_SPEED
                             (Sommerfeld)                                                   */
  int stop_rule;          /* NRPYELL_STOP_EVERY_STEP: the residual of the state before every
                             step is checked and the solve stops at log10 L2 <= log10_target
                             (max_steps caps the steps);
                             NRPYELL_STOP_ET_THORN: the Einstein Toolkit thorn's loop
                             (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:
                             181-223): with k = check_every, the checks after steps 1, 1+k,
                             1+2k, ... and after step max_steps+1 (max_steps = the thorn's
                             max_iterations) decide, the stop needs log10 L2 < log10_target,
                             and the solve ends after step max_steps+1 if none passes (status
                             NRPYELL_STATUS_MAX_STEPS). FP64 single-configuration solves only. */
  int substeps_per_launch;  /* FP64 steps of the resident solver: 2 = two RK4 substeps per
                             launch (two launches per step; radiation outer BC, Nxx2 >= 6); 1 =
                             one per launch; 0 = 2 where the grid and BC allow it, else 1. Both
                             give the same states bit for bit.                               */
} nrpyell_solve_options;
#define NRPYELL_STOP_EVERY_STEP 0
#define NRPYELL_STOP_ET_THORN 1

#define NRPYELL_OUTER_BC_RADIATION 0
#define NRPYELL_OUTER_BC_SOMMERFELD6 1
/* The wave speed the thorn's Sommerfeld BC uses on every grid
 * (conformally_flat_BBH_apply_bcs_sommerfeld.h:12): the outer-boundary wave speed of a
 * 64x64x4 grid with AMAX 1e6, SINHWAA 0.07, bScale 5. */
#define NRPYELL_ET_SOMMERFELD_SPEED 4.4793122676309114e+06

/* The Einstein Toolkit thorn's outer BC on the RHS, then the inner parity copy (device arrays).
 * Replaces conformally_flat_BBH_apply_bcs_sommerfeld (NRPyEllipticET/src/conformally_flat_BBH/
 * conformally_flat_BBH_apply_bcs_sommerfeld.h:139-222), which reads the thorn's own
 * bc_struct and fixes wavespeed = NRPYELL_ET_SOMMERFELD_SPEED, f_infinity = 0 and
 * radial_falloff_power = 3 (:10-12); here the tables are this ABI's bc_struct (the same ghost
 * maps) and the three are arguments. 7-point 6th-order derivatives; at each pure outer point
 *   rhs = -c (d_r f + (f - f_inf)/r) + (rhs_q + c (d_r f_q + (f_q - f_inf)/r_q)) (r_q/r)^n,
 * q the neighbour one point inward. NRPYELL_OK or NRPYELL_EINVAL. */
int nrpyell_apply_bcs_sommerfeld(const paramstruct *params, const bc_struct *bcstruct, REAL *xx[3],
                                 const REAL wavespeed[NUM_EVOL_GFS], const REAL f_infinity[NUM_EVOL_GFS],
                                 const REAL radial_falloff_power[NUM_EVOL_GFS], REAL *gfs, REAL *rhs_gfs);

typedef struct {
  int steps;              /* RK4 steps taken                                               */
  int fp32_steps;         /* of which in FP32                                              */
  double log10_l2;        /* last measured log10 L2 of the residual                         */
  double log10_l2_initial;/* log10 L2 of the residual of the initial guess                  */
  double seconds;         /* wall time of the solve (device-synchronised)                   */
  int converged;
  int status;             /* NRPYELL_STATUS_*: why the solve ended                          */
} nrpyell_solve_result;
#define NRPYELL_STATUS_MAX_STEPS 0  /* still relaxing when max_steps was reached            */
#define NRPYELL_STATUS_CONVERGED 1  /* log10 L2 <= log10_target                              */
#define NRPYELL_STATUS_NAN 2        /* the residual became NaN                               */
#define NRPYELL_STATUS_DIVERGED 3   /* the residual rose 4 decades above its lowest value    */

void nrpyell_default_options(nrpyell_solve_options *opt);
/* Build device tables for griddata (params, xx, bcstruct, rfmstruct, gridfuncs must be set
 * up as in the reference driver). The solver does not own griddata's arrays. */
int nrpyell_solver_create(griddata_struct *griddata, REAL dt, const nrpyell_solve_options *opt,
                          nrpyell_solver **out);
int nrpyell_solver_destroy(nrpyell_solver *s);
/* Relax y_n in place until log10 L2 <= target (or max_steps). */
int nrpyell_solver_solve(nrpyell_solver *s, nrpyell_solve_result *res);
/* Exactly n RK4 steps (graph replay when opt->steps_per_graph > 1), no residual checks. */
int nrpyell_solver_steps(nrpyell_solver *s, int n);
/* Fused residual + L2 of the current y_n; returns log10 L2 in *log10_l2. */
int nrpyell_solver_residual_l2(nrpyell_solver *s, double *log10_l2);
/* Time n launches of the fused substep kernel for stage `stage` (1..4; 0 = stage 1 with the
 * fused residual reduction; 5 and 6: the two-substep kernels of stages 1+2, with the residual
 * reduction, and 3+4) with HIP events on the solver's stream; mean microseconds. */
int nrpyell_solver_time_substep(nrpyell_solver *s, int stage, int n, double *us_per_launch);
/* Exactly n RK4 steps of the production FP64 path (graph replay, residual reduced every
 * step, no early stop), timed with HIP events on the solver's stream. */
int nrpyell_solver_timed_steps(nrpyell_solver *s, int n, double *ms, double *log10_l2);
/* A new damping strength mid-relaxation: params.eta_damping = eta and v += (eta - old eta) u at
 * every point, so that u and u_t = v - eta u carry over unchanged (the steady state, v = eta u,
 * moves with eta). Solvers built on griddata pick the new eta up at their next solve. Synchronous;
 * NRPYELL_OK or NRPYELL_EINVAL. */
int nrpyell_set_eta_damping(griddata_struct *gd, double eta);
/* RK4 substeps per launch of the solver's FP64 steps (1 or 2, see substeps_per_launch). */
int nrpyell_solver_substeps_per_launch(nrpyell_solver *s);
/* Residual history of the last solve: (steps completed, log10 L2) pairs; returns the count. */
int nrpyell_solver_history(nrpyell_solver *s, int max, int *steps, double *log10_l2);
/* Block placement of the solver's two-substep kernel: out[L] (L < max) = the block that the
 * L-th workgroup of a launch computes (BC blocks first, then interior blocks in tile order);
 * returns the launch's block count, or 0 when the kernel runs its own order (one-substep path,
 * or a launch that does not fit one round of resident workgroups). No counterpart in the
 * reference (a property of this GPU path): for tools that inspect the placement. */
int nrpyell_solver_block_order(const nrpyell_solver *s, unsigned short *out, int max);
/* The same placement as planned for a grid of N0 x N1 x N2 interior points with tiles of tile_i0
 * (16, 32, 64) x 256/tile_i0 points x `planes` (1, or 2 with tile_i0 >= 32) on a GPU of ncu CUs
 * in 8 XCDs, without a GPU: out[L] as above; returns the block count, 0 for the kernel's own
 * order, NRPYELL_EINVAL for invalid arguments. No counterpart in the reference. */
int nrpyell_block_order_plan(int ncu, int N0, int N1, int N2, int tile_i0, int planes, unsigned short *out, int max);

/* Batched FP64 solver: n configurations, each with its own griddata (any grid shape), relaxed
 * together -- one launch per RK4 substep per tile width covers every configuration, and each
 * configuration stops on its own residual. tile_i0[k] and planes[k] (see
 * nrpyell_solve_options; 0 or a NULL array for the default) are the tile chosen for
 * configuration k (by the autotuner); configurations with the same tile share a launch. */
typedef struct nrpyell_batch nrpyell_batch;
int nrpyell_batch_create(griddata_struct **griddatas, const double *dts, const int *tile_i0, const int *planes,
                         int n, const nrpyell_solve_options *opt, nrpyell_batch **out);
/* Relax every configuration to opt->log10_target; res has n entries. */
int nrpyell_batch_solve(nrpyell_batch *b, nrpyell_solve_result *res, double *seconds);
int nrpyell_batch_destroy(nrpyell_batch *b);

/* Learned initial guess: coordinate MLP (64 padded input features -> 5 x [Linear 256 + SiLU]
 * -> Linear 1) on BF16 MFMA kernels with FP32 accumulation; FP32 master weights and Adam.
 * Parameter layout (FP32): for l = 1..5 W_l [256 x K_l] (K_1 = 64, else 256), b_l [256];
 * then w_out [256], b_out [1]. Row counts must be multiples of 64. */
typedef struct nrpyell_mlp nrpyell_mlp;
int nrpyell_mlp_create(unsigned long long seed, nrpyell_mlp **out);
int nrpyell_mlp_destroy(nrpyell_mlp *m);
long nrpyell_mlp_num_params(nrpyell_mlp *m);
int nrpyell_mlp_get_params(nrpyell_mlp *m, float *host);
int nrpyell_mlp_set_params(nrpyell_mlp *m, const float *host);
/* x: device BF16 [M x 64]; y: device FP32 [M] */
int nrpyell_mlp_forward(nrpyell_mlp *m, const void *x, int M, float *y);
/* mean squared error against device targets t [M], gradients (host, num_params); the weight
 * gradients are reduced over fixed row ranges in a fixed order: repeated calls agree bit for bit */
int nrpyell_mlp_loss_grad(nrpyell_mlp *m, const void *x, const float *t, int M, double *loss, float *grads_host);
/* one Adam step with the gradient of the last nrpyell_mlp_loss_grad call */
int nrpyell_mlp_adam_step(nrpyell_mlp *m, float lr);
/* nsteps Adam steps on uniform minibatches from a device dataset (x [N x 64], t [N]); three
 * launches per step (forward + loss + data-gradient chain, weight gradients, Adam); deterministic */
int nrpyell_mlp_train(nrpyell_mlp *m, const void *x, const float *t, long N, int batch, int nsteps, float lr,
                      unsigned long long seed, double *losses_host);
/* rows (interior + pure outer points, rounded up to 64) of a grid's feature matrix */
long nrpyell_mlp_feature_rows(const griddata_struct *gd);
/* device features x [rows x 64] (BF16), per-row u scale and grid index of a griddata */
int nrpyell_mlp_features(const griddata_struct *gd, void *x, double *scale, long *idx);
/* device training targets t = log(u/scale) from griddata's y_n at the feature rows */
int nrpyell_mlp_targets(const griddata_struct *gd, const double *scale, const long *idx, float *t);
/* y_n <- network guess: u at interior and pure outer points, smoothed by smooth_passes
 * binomial passes along i0, i1, i2 (0..64; k_odd_gfs is scratch), v = eta u, parity images */
int nrpyell_mlp_seed(nrpyell_mlp *m, griddata_struct *gd, int smooth_passes);
/* n_take uniform random rows of a device (x, t) set copied into (x_dst, t_dst) */
int nrpyell_mlp_sample_rows(const void *x_src, const float *t_src, long n_src, long n_take, unsigned long long seed,
                            void *x_dst, float *t_dst);
int nrpyell_memcpy_d2d(void *dst, const void *src, size_t bytes);

/* ---- Einstein Toolkit consumer path (NRPyEllipticET/src/conformally_flat_BBH/) ---------- */

/* Options of the thorn's Initialize_ADMBase (NRPyEllipticET/param.ccl:39-88,142-145). */
#define NRPYELL_ORBITAL_PLANE_XY 0 /* "xy"/"yx" (default): (x,y,z)_NRPy = (y,z,x)_ETK */
#define NRPYELL_ORBITAL_PLANE_XZ 1 /* "xz"/"zx": same axes                             */
#define NRPYELL_LAPSE_PSI_N 0      /* "NRPyEllipticET-psi^n": alpha = psi^n             */
#define NRPYELL_LAPSE_AVERAGED 1   /* "NRPyEllipticET-averaged"                         */
#define NRPYELL_ADM_NOUT 17
typedef struct {
  int orbital_plane;
  REAL position_shift[3];
  int initial_lapse;
  REAL lapse_exponent_n;   /* param.ccl default -2            */
  REAL NRPy_epsilon;       /* param.ccl default 1e-8          */
  int interpolation_order; /* Lagrange order 1..6, default 4  */
} nrpyell_admbase_options;

/* Replaces NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase
 * (NRPyEllipticET_conformally_flat_BBH_Initialize_ADMBase.c:5-26), with its interpolation call
 * (..._interpolate_solution_to_ADMBase.c:4-61) and ADM formulas
 * (conformally_flat_BBH_compute_ADM_Cartesian_quantities_from_uu.c:7-103). griddata holds the
 * solved u (y_n_gfs, ghosts filled, as every nrpyell solve leaves it), the SinhSymTP xx (host)
 * and the NRPy-frame parameters the solve ran with. cctk_lsh[0]*[1]*[2] points; x, y, z and
 * the 17 outputs are device arrays. Synchronous. NRPYELL_ENAN when a point's coordinates stay
 * NaN after the NRPy_epsilon retry (the thorn aborts there), NRPYELL_EINVAL for bad options or
 * a point outside the grid. */
int nrpyell_Initialize_ADMBase(const griddata_struct *griddata, const nrpyell_admbase_options *opt,
                               const int *cctk_lsh, const REAL *x, const REAL *y, const REAL *z, REAL *uuGF,
                               REAL *alp, REAL *betax, REAL *betay, REAL *betaz, REAL *gxx, REAL *gxy, REAL *gxz,
                               REAL *gyy, REAL *gyz, REAL *gzz, REAL *kxx, REAL *kxy, REAL *kxz, REAL *kyy,
                               REAL *kyz, REAL *kzz);
/* conformally_flat_BBH_compute_TRC (conformally_flat_BBH_set_optimum_eta_damping.c:3-16) */
REAL compute_TRC(const REAL A, const REAL b, const REAL C_0, const REAL dt, const REAL dxx1, const REAL dxx2);
/* conformally_flat_BBH_set_optimum_eta_damping (conformally_flat_BBH_set_optimum_eta_damping.c:18-27):
 * params->eta_damping = 6.9 TRC^-0.79 - 0.7 from AMAX, bScale, dxx1, dxx2 */
void set_optimum_eta_damping(const REAL dt, const REAL CFL_FACTOR, paramstruct *params);

#ifdef __cplusplus
}
#endif
#endif
