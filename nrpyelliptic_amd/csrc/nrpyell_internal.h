// nrpyell_internal.h -- declarations shared by the translation units of libnrpyell_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "nrpyelliptic_amd.h"
#include "nrpyell_common.h"

#define NRPYELL_HIST 4096

// Device-resident convergence state of one solve.
struct SolverDeviceState {
  int done;          // 0 running, 1 converged, 2 NaN, 3 diverged, 4 step cap of the thorn's rule
  int steps;         // RK4 steps completed at the last residual check
  int nhist;
  int nchecks;       // residual checks made by fused stage-1 kernels
  double last_log10;
  int have_best;
  double best_log10;  // lowest residual seen
  // stopping rule: rule_every == 0: every check decides, stop at log10 L2 <= target (the
  // resident solver's rule); rule_every = k >= 1: the Einstein Toolkit thorn's rule -- only the
  // checks after steps 1, 1+k, 1+2k, ... and after step rule_last decide, the stop needs
  // log10 L2 < target, and the solve ends after step rule_last in any case
  int rule_every;
  int rule_last;
  int hist_step[NRPYELL_HIST];
  double hist_log10[NRPYELL_HIST];
};

// A residual this many decades above the lowest one seen ends the solve as diverged: the
// relaxation diverges for some puncture parameters (the reference's C as well), and a batch
// should not spend max_steps on them. Transient rises of converging solves stay below one
// decade.
#define NRPYELL_DIVERGE_DECADES 4.0

// threads per block of the fused substep kernel (128, 256 or 512; 256 is the fastest at
// 128x128x16), and the i1 rows handled by one of its BC blocks
#ifndef NRPYELL_BLOCK
#define NRPYELL_BLOCK 256
#endif
#ifndef NRPYELL_BC_ROWS
#define NRPYELL_BC_ROWS (NRPYELL_BLOCK / 16)
#endif
// i1 rows per BC block of the two-substep kernel (each block also works on 10 halo rows; 8 and
// 16 measured alike at 128x128x16 with 32x2 tiles, 16 faster with one-plane tiles)
#ifndef NRPYELL_PAIR_BC_ROWS
#define NRPYELL_PAIR_BC_ROWS 16
#endif

// Bookkeeping of one residual check (log10 L2 = l) taken after st->steps steps: history,
// done = 1 at the target, 2 on NaN, 3 on divergence.
__device__ inline void record_check(SolverDeviceState *st, double l, double target) {
  st->last_log10 = l;
  if (st->nhist < NRPYELL_HIST) {
    st->hist_step[st->nhist] = st->steps;
    st->hist_log10[st->nhist] = l;
    st->nhist++;
  }
  if (!(l == l)) {
    st->done = 2;
    return;
  }
  if (!st->have_best || l < st->best_log10) {
    st->best_log10 = l;
    st->have_best = 1;
  }
  if (st->rule_every > 0) {
    // NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:185-223: for n = 0..N_final,
    // step, then check when n % k == 0 or n == N_final (the state after n + 1 steps), stop on
    // log10 L2 < target
    const int m = st->steps;
    const bool due = m >= 1 && ((m - 1) % st->rule_every == 0 || m == st->rule_last);
    if (due && l < target) st->done = 1;
    else if (m >= st->rule_last) st->done = 4;
    else if (l > st->best_log10 + NRPYELL_DIVERGE_DECADES) st->done = 3;
    return;
  }
  if (l <= target) st->done = 1;
  else if (l > st->best_log10 + NRPYELL_DIVERGE_DECADES) st->done = 3;
}

struct Punctures {
  double m[2];
  double pos[2][3];
  double P[2][3];
  double S[2][3];
};

// The Bowen-York conformal extrinsic curvature at the Cartesian point X, and h_n = m_n/(2 r_n)
// (psi_bg = 1 + h_0 + h_1). Per puncture, with x = X - x_n and r = |x|:
//   A_ij = (3/2)(P_i x_j + P_j x_i)/r^3 - (3/2)(x.P) delta_ij/r^3 + (3/2)(x.P) x_i x_j/r^5
//          - 3[(x cross S)_i x_j + (x cross S)_j x_i]/r^5
// (NRPyElliptic_SourceTerm.py:48-143 differentiated by hand).
__host__ __device__ inline void bowen_york(const Punctures &pu, const double X[3], double A[3][3], double h[2]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] = 0.0;
  for (int n = 0; n < 2; n++) {
    const double x[3] = {X[0] - pu.pos[n][0], X[1] - pu.pos[n][1], X[2] - pu.pos[n][2]};
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], r = sqrt(r2);
    const double ir3 = 1.0 / (r2 * r), ir5 = ir3 / r2;
    const double *P = pu.P[n], *S = pu.S[n];
    const double xP = x[0] * P[0] + x[1] * P[1] + x[2] * P[2];
    const double xS[3] = {x[1] * S[2] - x[2] * S[1], x[2] * S[0] - x[0] * S[2], x[0] * S[1] - x[1] * S[0]};
    h[n] = pu.m[n] / (2.0 * r);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        A[i][j] += 1.5 * (P[i] * x[j] + P[j] * x[i]) * ir3 - (i == j ? 1.5 * xP * ir3 : 0.0) +
                   1.5 * xP * x[i] * x[j] * ir5 - 3.0 * (xS[i] * x[j] + xS[j] * x[i]) * ir5;
  }
}

// One outer point for the drop-in radiation BC: point p, interior neighbour q,
// clamped stencil offsets along x0,x1,x2 for each, geometry r, dxx0/dr, dxx1/dr, dxx2/dr.
struct RadPoint {
  long p, q;
  int off[2][3];
  double geo[2][4];
};
// Wave speed, value at infinity and radial falloff power per evolved gridfunction of the
// Einstein Toolkit thorn's Sommerfeld BC.
struct SommerfeldParams {
  double c[2], finf[2], falloff[2];
};
struct RadLayers {
  const RadPoint *l[NRPYELL_NG];  // the x0-face list of each ghost layer, same row order
};

// Arguments of the fused RK4 substep kernel.
template <typename T>
struct SubstepArgs {
  GridDims g;
  const T *in;    // stage input (u | v)
  T *out;         // stage output (stage 1-3) or y_n (stage 4)
  const T *yn;    // y_n
  T *rt;          // running total
  const T *src;   // per point (psi, 0.125 c^2 AA^2), c the wave speed
  const T *coef;  // per (i0,i1): c^2 x the 5 metric coefficients, then the L2 weight / c^4
  const T *geo;   // outer-face geometry: [pt 0..5][i1][1/r, dxx0/dr, dxx1/dr, (r_0/r_pt)^3], pt 0 = last interior i0
  T cOB_u, cOB_v;   // wave speed of the outer BC
  int bc_kind;      // NRPYELL_OUTER_BC_RADIATION or NRPYELL_OUTER_BC_SOMMERFELD6
  T eta, dt;
  const int *done;  // nonzero: converged, every launch returns at once
  int tx;           // tile width along i0 (threads), tile height along i1 = blockDim/tx
  int pl;           // i2 planes per thread (1 or 2)
  int nbc_rows;     // BC blocks per i2 plane: ceil(N1 / NRPYELL_BC_ROWS)
  int ntile0, ntile1;
  int i0_first;     // two-substep kernel: i0 of the first tile's column 0 (tiles end at ilast)
  // div_mul of N2/pl, ntile0 and nbc_rows: the kernels split block indices by them
  unsigned long long ng_mul, nt0_mul, nbc_mul;
  // Stage 1 (FP64) can also reduce the residual of its input y_n: the stencil it evaluates
  // is c^2 times the residual (weighted by the L2 weight / c^4 of the coefficient record).
  // Non-null l2_partial enables it; block 0 of the stage-2 launch that follows sums the
  // partials and updates the convergence state.
  double *l2_partial;         // one partial sum per block
  SolverDeviceState *state;   // convergence state
  double l2_volume, l2_target;
};


// kernel launchers (nrpyell_kernels.hip)
void launch_rhs_exact(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *in, double *out,
                      double eta, hipStream_t st);
void launch_residual_exact(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *in,
                           double *out, hipStream_t st);
// the solver's tables from rfm and auxevol: metric coefficients (coef64), their c^2-folded record
// with the L2 weight (kc64, kc32), the per-point source (src); *bad raised where the wave speed
// varies along xx2; *bad and *also_clear (when given) zeroed first, by the same launch
void launch_tables(const GridDims &g, const rfm_struct *rfm, const double *aux, const double *w2d, double *coef64,
                   double *kc64, float *kc32, double *src, int *bad, int *also_clear, hipStream_t st);
template <typename T>
void launch_fill_ghost_planes(const GridDims &g, T *y, int ngf, hipStream_t st);
// dst = src with dst's ghost phi planes filled from their sources (copy + fill_ghost_planes)
void launch_settle_fill(const GridDims &g, const double *src, double *dst, int ngf, hipStream_t st);
template <typename T>
void launch_substep(const SubstepArgs<T> &a, int stage, hipStream_t st);
// two RK4 substeps (first_stage 1: stages 1+2; 3: stages 3+4) in one launch (nrpyell_substep2.hip)
// Placement of the two-substep kernel's blocks for a launch that fits the GPU in one round:
// workgroup L runs as logical block q[L] (< nbc_rows * N2: a BC block, else interior block q[L] -
// that). Built on the host (pair_block_order) so that the CUs' loads even out; n == 0: the
// kernel's own order (xcd_logical). Passed by value in the kernel arguments, so a block reads its
// entry with the other arguments.
#define NRPYELL_ORDER_MAX 1024
struct BlockOrder {
  int n;
  unsigned q2[NRPYELL_ORDER_MAX / 2];  // entry L in bits 16 (L & 1) .. 16 (L & 1) + 15 of q2[L / 2]
  __host__ __device__ unsigned at(unsigned L) const {
    const unsigned w = q2[(L < NRPYELL_ORDER_MAX ? L : NRPYELL_ORDER_MAX - 1) >> 1];
    return (L & 1) ? w >> 16 : w & 0xffffu;
  }
  void set(unsigned L, unsigned v) {
    unsigned &w = q2[L >> 1];
    w = (L & 1) ? (w & 0xffffu) | (v << 16) : (w & 0xffff0000u) | v;
  }
};
template <typename T>
void launch_substep_pair(const SubstepArgs<T> &a, int first_stage, hipStream_t st, const BlockOrder *ord = nullptr);
// the same for many configurations in one launch (block maps as launch_substep_batched)
void launch_substep_pair_batched(const SubstepArgs<double> *cfgs, const int *block_cfg, const int *block_first,
                                 int nblocks, int tx, int pl, int first_stage, hipStream_t st);
template <typename T>
void launch_substep_batched(const SubstepArgs<T> *cfgs, const int *block_cfg, const int *block_first, int nblocks,
                            int tx, int pl, int stage, hipStream_t st);
long launch_residual_l2(const GridDims &g, const void *u, bool u_is_float, const double *aux, const double *coef,
                        const double *w2d, double *partial, const int *done, double *res_out, hipStream_t st);
long launch_l2_partial(const GridDims &g, const double *gf, const double *w2d, double *partial, hipStream_t st);
long l2_partial_blocks(const GridDims &g);  // partial sums launch_l2_partial writes
// Grid constants of the L2 volume weights computed on the device.
struct L2WeightParams {
  double AMAX, SINHWAA, bScale, dxx0, dxx1, dxx2, radius;
};
void launch_l2_weights(const GridDims &g, const L2WeightParams &q, const double *xx0, const double *xx1, double *w2d,
                       hipStream_t st);
void launch_l2_finalize_weights(const double *partial, long npart, const double *w2d, long nw, int n2,
                                double *log10_out, hipStream_t st);
void launch_l2_finalize(const double *partial, long npart, double volume, SolverDeviceState *dst, int steps_per_check,
                        double target, double *log10_out, hipStream_t st);
void launch_inner_gather(long nt, int n, const int *dst, const int *src, const int8_t *par0, double *gfs,
                         hipStream_t st);
void launch_radiation_rows(const GridDims &g, int n, const RadLayers &L, const double *gfs, double *rhs, double c_u,
                           double c_v, double finf_u, double finf_v, hipStream_t st);
void launch_sommerfeld_list(const GridDims &g, int n, const RadPoint *pts, const double *gfs, double *rhs,
                            const SommerfeldParams &sp, hipStream_t st);
void launch_radiation_list(const GridDims &g, int n, const RadPoint *pts, const double *gfs, double *rhs, double c_u,
                           double c_v, double finf_u, double finf_v, hipStream_t st);
void launch_rfm(const GridDims &g, double AMAX, double w, double b, const double *xx0_d, const double *xx1_d,
                const rfm_struct &r, hipStream_t st);
void launch_auxevol(const GridDims &g, double AMAX, double w, double b, const Punctures &pu, const double *xx0_d,
                    const double *xx1_d, const double *xx2_d, double *aux, hipStream_t st);
void launch_wavespeed(const GridDims &g, const double *ds2d, double CFL, double dt, double *aux, hipStream_t st);
void launch_convert_d2f(long n, const double *in, float *out, hipStream_t st);
void launch_convert_f2d(long n, const float *in, double *out, hipStream_t st);
void launch_fill(long n, double v, double *out, hipStream_t st);
void launch_shift_v(long n, double d, double *y, hipStream_t st);

// host setup (nrpyell_host.cpp)
GridDims dims_from_params(const paramstruct *p);
void host_min_ds_2d(const paramstruct *p, REAL *const xx[3], double *ds2d);  // G0*G1, interior entries
void host_outer_geometry(const paramstruct *p, REAL *const xx[3], double *geo);  // 6*G1*4
void host_l2_weights(const paramstruct *p, REAL *const xx[3], double radius, double *w2d, double *volume);
void host_radial_geometry(const paramstruct *p, double x0, double x1, double out[4]);
int host_bcstruct(const paramstruct *params, REAL *xx[3], bc_struct *bc, std::string *err);
// checks the closed-form ghost map against a bcstruct table; returns 0 when identical
int ghostmap_verify(const GridDims &g, const bc_struct *bc, std::string *err);

// error plumbing (nrpyell_capi.hip)
void set_last_error(const std::string &s);
[[noreturn]] void fatal(const char *what, const std::string &msg);
// A failed HIP runtime call inside the library. The nrpyell_* entry points turn it into
// NRPYELL_EHIP with the message in nrpyell_last_error() (api_call); the reference-surface entry
// points, void like the reference's, print it and exit(1) as the reference does (ref_call).
struct HipFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw HipFailure(std::string(#x) + " failed: " + hipGetErrorString(e_) + " at " + __FILE__ + \
                       ":" + std::to_string(__LINE__));                                            \
  } while (0)

// Device buffer owned by one call: freed on every return path, HipFailure included.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  explicit DevBuf(size_t n) { HIPCHK(hipMalloc((void **)&p, (n > 0 ? n : 1) * sizeof(T))); }
  ~DevBuf() { (void)hipFree(p); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

template <typename F>
int api_call(F &&f) noexcept {
  try {
    return f();
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return NRPYELL_EHIP;
  }
}

template <typename F>
auto ref_call(const char *name, F &&f) noexcept -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception &e) {
    fatal(name, e.what());
  }
}
