// nrpyell_substep.hip -- the fused RK4 substep kernel (the hot path).
//
// One launch does, for one RK4 substep (MoL_step_forward_in_time.c:11-176):
//   1. RHS at every interior point: 10th-order star stencil of u (rhs_eval.c:86-110) with the
//      precombined (i0,i1) metric coefficients, v - eta u, and c^2 (lap u + AA^2/(8 (psi+u)^7));
//   2. the radiation condition on the 5 outer ghost layers of the x0-max face (the reference's
//      innermost-first recurrence, apply_bcs_outerradiation_and_inner.c:369-408, in closed
//      form), done by separate light BC blocks in the same launch, from LDS;
//   3. the stage update (RK4 k_odd/k_even scheme) of every interior and outer point;
//   4. the inner ghost zones: each updated point writes its own parity images (the closed form
//      of the bcstruct map, verified against the table on the host), which is the inner-BC
//      gather of apply_bcs_inner_only.c:18-26 applied to the stage output (parity +1 commutes
//      with the linear stage update, so the values are those of the reference ordering).
//      Only the images inside the interior phi planes are written: phi neighbours are read
//      through the periodic map, so the ghost planes are dead during a solve and are filled
//      once when the state is handed back (fill_ghost_planes_kernel);
//   5. (stage 1, FP64) the L2 of the residual of the stage input y_n: the stencil of step 1 is
//      c^2 times the residual, so a block reduction and a last-block finalize give the
//      reference's every-step convergence check (main.c:146-151) for free.
//
// Interior tiles are TX (i0, across lanes) x TY (i1) points of one or two i2 planes,
// NRPYELL_BLOCK threads; the u plane with its i0/i1 halos is staged in LDS, the i2 arm is read from global
// memory (coalesced).
#include <hip/hip_runtime.h>
#include <math.h>

#include "nrpyell_common.h"
#include "nrpyell_internal.h"
#include "nrpyell_stencil.h"

#define NG NRPYELL_NG

// ---------------------------------------------------------------------------------------
// Interior blocks: tile `tile` (t0 + ntile0*t1) of the PL consecutive i2 planes
// NG + PL*pgroup ...; `bid` is the block's index among the nblk interior blocks of this
// configuration (for the L2 partial sums). With PL = 2 every thread updates the same (i0,i1)
// on two planes: the metric coefficients and weights are read once, and the i2 arms share 12
// planes instead of 2 x 11. The outer layers belong to the BC blocks (bc_body).
// ---------------------------------------------------------------------------------------
template <typename T, int STAGE, int TX, int PL>
__device__ __forceinline__ void interior_body(const SubstepArgs<T> &a, const int tile, const int pgroup,
                                              const unsigned bid, const unsigned nblk) {
  constexpr int NT = NRPYELL_BLOCK, TY = NT / TX;
  constexpr int BX = TX + 2 * NG, BY = TY + 2 * NG, BOX = BX * BY, NR = (BOX + NT - 1) / NT;
  constexpr int NW = PL + 2 * NG;  // i2 window: planes i2a-NG .. i2a+PL-1+NG
  __shared__ T su[PL][NR * NT];    // u on this block's planes: tile + halos (padded: no store guards)
  __shared__ double s_red[NT / 64];

  const GridDims &g = a.g;
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int t1 = (int)div_by((unsigned)tile, a.nt0_mul), t0 = tile - t1 * a.ntile0;
  const int i2a = NG + PL * pgroup;
  const int i0s = NG + t0 * TX, i1s = NG + t1 * TY;
  const int i0 = i0s + tx, i1 = i1s + ty;
  const bool valid = i0 < g.N0 + NG && i1 < g.N1 + NG;
  // 32-bit element indices (the host caps a configuration at 2^32 bytes per array)
  const unsigned nt = (unsigned)g.ntot, s1 = g.G0, s2 = (unsigned)(g.G0 * g.G1);
  const bool fuse_l2 = (STAGE == 1) && (a.l2_partial != nullptr);
  const int done = *a.done;

  // ---- global loads. Every load is unconditional, from an index clamped into the grid:
  // values the stencils do not reach (box corners, columns or rows past a ragged edge) are
  // loaded but never read, and the loads compile to straight-line code the compiler keeps in
  // flight across the per-plane barriers. Issue order is plane-major (vmcnt counts in order),
  // so plane 0 computes and stores while plane 1's loads are in flight.
  const int i0c = min(i0, g.N0 + NG - 1), i1c = min(i1, g.N1 + NG - 1);
  const unsigned k2 = (unsigned)i1c * s1 + i0c;
  T stage_u[PL][NR];
  T zw[NW];  // u on the i2 window at (i0,i1), through the periodic map
  T vv[PL], psi[PL], AAc[PL], yn_u[PL], yn_v[PL], rt_u[PL], rt_v[PL], K[5];
  double w2 = 0.0;
#pragma unroll
  for (int q = 0; q < PL; q++) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int k = tid + NT * r;
      const int bx = k % BX, by = k / BX;
      const int j0 = min(i0s - NG + bx, g.G0 - 1), j1 = min(i1s - NG + by, g.G1 - 1);
      stage_u[q][r] = ld32(a.in, (unsigned)(i2a + q) * s2 + (unsigned)j1 * s1 + j0);
    }
    if (q == 0) {
      // the window spans plane indices i2a - NG .. i2a + PL - 1 + NG: within one period of the
      // interior planes once N2 >= 6
      if (g.N2 >= 6) {
#pragma unroll
        for (int jj = 0; jj < NW; jj++) zw[jj] = ld32(a.in, (unsigned)wrap2_near(g, i2a - NG + jj) * s2 + k2);
      } else {
#pragma unroll
        for (int jj = 0; jj < NW; jj++) zw[jj] = ld32(a.in, (unsigned)wrap2(g, i2a - NG + jj) * s2 + k2);
      }
      T kw;
      ld2x32(a.coef, 6 * k2, K[0], K[1]);
      ld2x32(a.coef, 6 * k2 + 2, K[2], K[3]);
      ld2x32(a.coef, 6 * k2 + 4, K[4], kw);
      if (fuse_l2) w2 = (double)kw;
    }
    const unsigned p = (unsigned)(i2a + q) * s2 + k2;
    vv[q] = ld32(a.in, nt + p);
    ld2x32(a.src, 2 * p, psi[q], AAc[q]);
    yn_u[q] = 0; yn_v[q] = 0; rt_u[q] = 0; rt_v[q] = 0;
    if (STAGE != 1) {  // stage 1 updates its own input (y_n)
      yn_u[q] = ld32(a.yn, p);
      yn_v[q] = ld32(a.yn, nt + p);
      rt_u[q] = ld32((const T *)a.rt, p);
      rt_v[q] = ld32((const T *)a.rt, nt + p);
    }
  }
  if (done) return;  // converged: uniform across the grid

  double l2acc = 0.0;
  const int cb = (ty + NG) * BX + tx + NG;
#pragma unroll
  for (int q = 0; q < PL; q++) {
#pragma unroll
    for (int r = 0; r < NR; r++) su[q][tid + NT * r] = stage_u[q][r];
    __syncthreads();
    if (valid) {
      const int i2 = i2a + q;
      const unsigned p = (unsigned)i2 * s2 + k2;
      const T uu = zw[NG + q];
      T xm[6], xp[6], ym[6], yp[6], zm[6], zp[6];
#pragma unroll
      for (int k = 1; k <= 5; k++) {
        xm[k] = su[q][cb - k]; xp[k] = su[q][cb + k];
        ym[k] = su[q][cb - k * BX]; yp[k] = su[q][cb + k * BX];
        zm[k] = zw[NG + q - k]; zp[k] = zw[NG + q + k];
      }
      if (STAGE == 1) {
        yn_u[q] = uu;
        yn_v[q] = vv[q];
      }
      T ku, kv;
      point_rhs<T>(uu, vv[q], psi[q], AAc[q], K, xm, xp, ym, yp, zm, zp, a.eta, ku, kv);
      if (fuse_l2) l2acc += w2 * (double)kv * (double)kv;  // w / c^4 x (c^2 R)^2
      T o_u, o_v;
      stage_update<T, STAGE>(ku, a.dt, yn_u[q], rt_u[q], o_u);
      stage_update<T, STAGE>(kv, a.dt, yn_v[q], rt_v[q], o_v);
      if (STAGE != 4) {
        st32(a.rt, p, rt_u[q]);
        st32(a.rt, nt + p, rt_v[q]);
      }
      st32(a.out, p, o_u);
      st32(a.out, nt + p, o_v);
      for_each_image_in_planes(g, i0, i1, i2, [&](long d) {
        st32(a.out, (unsigned)d, o_u);
        st32(a.out, nt + (unsigned)d, o_v);
      });
    }
  }

  // ---- residual L2 of the stage-1 input: one partial per block (plain store); the stage-2
  // launch that follows reduces them in block order, after the kernel boundary has made them
  // visible (no atomics, no L2 flushes inside the kernel).
  if (fuse_l2) {
    for (int o = 32; o > 0; o >>= 1) l2acc += __shfl_down(l2acc, o, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = l2acc;
    __syncthreads();
    if (tid == 0) a.l2_partial[bid] = wave_sum<NT / 64>(s_red);
  }
  if (STAGE == 2 && a.l2_partial != nullptr && bid == 0) {
    double acc = 0.0;
    for (unsigned i = tid; i < nblk; i += NT) acc += a.l2_partial[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    __syncthreads();
    if ((tid & 63) == 0) s_red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
      const double S = wave_sum<NT / 64>(s_red);
      const double l = log10(1e-16 + sqrt(S / a.l2_volume));
      SolverDeviceState *st = a.state;
      st->steps = st->nchecks;  // completed steps when stage 1 read y_n
      st->nchecks++;
      // once done is set, stages 3 and 4 of this step skip, leaving y_n as it was after
      // st->steps steps
      record_check(st, l, a.l2_target);
    }
  }
}

// ---------------------------------------------------------------------------------------
// BC blocks: the radiation condition on the 5 outer layers (i0 = ilast+1..ilast+5) of
// NRPYELL_BC_ROWS rows (i1) of one i2 plane. Each row needs the RHS at the last interior point
// (re-evaluated here with point_rhs), and u and v on the 12 columns ilast-6..ilast+5 with the
// rows' i1 halos for the one-sided i0 derivatives and the centred i1 derivatives
// (apply_bcs_outerradiation_and_inner.c:11-141). With the radiation term A_pt = -c (df/dr + f/r)
// at layer pt (pt = 0: the last interior point), the reference's innermost-first recurrence
// (:369-408)
//   rhs_pt = A_pt + (r_{pt-1}/r_pt)^3 (rhs_{pt-1} - A_{pt-1})
// telescopes to rhs_pt = A_pt + (r_0/r_pt)^3 (rhs_0 - A_0), so the layers are independent
// (1/r and (r_0/r_pt)^3 come precomputed in the geometry table). The blocks are light and are
// dispatched first; the interior blocks carry no BC work.
// ---------------------------------------------------------------------------------------
template <typename T, int STAGE>
__device__ __forceinline__ void bc_body(const SubstepArgs<T> &a, const int rb, const int plane) {
  constexpr int TYB = NRPYELL_BC_ROWS, BYB = TYB + 2 * NG, NBX = 12, NBOX = NBX * BYB;
  constexpr int NT = NRPYELL_BLOCK, NRB = (NBOX + NT - 1) / NT, NGEO = (24 * TYB + NT - 1) / NT;
  static_assert(12 * TYB <= NT && 10 * TYB <= NT, "one BC work item per thread");
  __shared__ T bu[NRB * NT], bv[NRB * NT];  // [by][c]: u, v at column ilast-6+c, row i1s-NG+by
  __shared__ T s_rhs0[TYB][2], s_A[TYB][2][6], s_geo[6][TYB][4], s_fd[NRPYELL_FD_TABLE];

  const GridDims &g = a.g;
  const int tid = threadIdx.x;
  const unsigned nt = (unsigned)g.ntot, s1 = g.G0, s2 = (unsigned)(g.G0 * g.G1);
  const int ilast = g.N0 + NG - 1;
  const int i1s = NG + rb * TYB, i2 = NG + plane;
  const unsigned pl = (unsigned)i2 * s2;
  const int done = *a.done;

  T st_u[NRB], st_v[NRB];
#pragma unroll
  for (int r = 0; r < NRB; r++) {
    const int k = tid + NT * r, c = k % NBX, by = min(k / NBX, BYB - 1);
    const unsigned q = pl + (unsigned)min(i1s - NG + by, g.G1 - 1) * s1 + max(ilast - NG - 1 + c, 0);
    st_u[r] = ld32(a.in, q);
    st_v[r] = ld32(a.in, nt + q);
  }
  T geo_r[NGEO];
#pragma unroll
  for (int r = 0; r < NGEO; r++) {
    const int k = min(tid + NT * r, 24 * TYB - 1);
    const int pt = k / (4 * TYB), row = (k / 4) % TYB, comp = k % 4;
    geo_r[r] = ld32(a.geo, ((unsigned)pt * g.G1 + min(i1s + row, g.G1 - 1)) * 4 + comp);
  }
  // RHS inputs at the last interior point of row tid (threads tid < TYB)
  const int rrow = min(tid, TYB - 1);
  const unsigned k2 = (unsigned)min(i1s + rrow, g.N1 + NG - 1) * s1 + ilast;
  T zw[2 * NG + 1], K[5], vv = 0, psi = 1, AAc = 0;
  if (tid < TYB) {
    if (g.N2 >= 6) {  // the window is within one period of the interior planes
#pragma unroll
      for (int j = 0; j <= 2 * NG; j++)
        zw[j] = j == NG ? (T)0 : ld32(a.in, (unsigned)wrap2_near(g, i2 - NG + j) * s2 + k2);
    } else {
#pragma unroll
      for (int j = 0; j <= 2 * NG; j++) zw[j] = j == NG ? (T)0 : ld32(a.in, (unsigned)wrap2(g, i2 - NG + j) * s2 + k2);
    }
    T kw;
    ld2x32(a.coef, 6 * k2, K[0], K[1]);
    ld2x32(a.coef, 6 * k2 + 2, K[2], K[3]);
    ld2x32(a.coef, 6 * k2 + 4, K[4], kw);
    vv = ld32(a.in, nt + pl + k2);
    ld2x32(a.src, 2 * (pl + k2), psi, AAc);
  }
  // outer-layer point of this thread (work item = (row, gf, layer 1..5))
  const int od_row = min(tid / 10, TYB - 1), od_gf = (tid % 10) / 5, od_pt = tid % 5 + 1;
  const unsigned od_p = pl + (unsigned)min(i1s + od_row, g.G1 - 1) * s1 + (ilast + od_pt) + od_gf * nt;
  const bool od_on = tid < 10 * TYB && i1s + od_row < g.N1 + NG;
  const T od_yn = ld32(a.yn, od_p);
  const T od_rt = STAGE != 1 ? ld32((const T *)a.rt, od_p) : (T)0;
  const T fdv = (T)load_fd_table(tid);
  if (done) return;
#pragma unroll
  for (int r = 0; r < NRB; r++) {
    bu[tid + NT * r] = st_u[r];
    bv[tid + NT * r] = st_v[r];
  }
#pragma unroll
  for (int r = 0; r < NGEO; r++)
    if (tid + NT * r < 24 * TYB) (&s_geo[0][0][0])[tid + NT * r] = geo_r[r];
  if (tid < NRPYELL_FD_TABLE) s_fd[tid] = fdv;
  __syncthreads();

  if (tid < TYB && i1s + tid < g.N1 + NG) {
    const int cb = (tid + NG) * NBX + NG + 1;  // column ilast
    T xm[6], xp[6], ym[6], yp[6], zm[6], zp[6];
#pragma unroll
    for (int k = 1; k <= 5; k++) {
      xm[k] = bu[cb - k]; xp[k] = bu[cb + k];
      ym[k] = bu[cb - k * NBX]; yp[k] = bu[cb + k * NBX];
      zm[k] = zw[NG - k]; zp[k] = zw[NG + k];
    }
    T ku, kv;
    point_rhs<T>(bu[cb], vv, psi, AAc, K, xm, xp, ym, yp, zm, zp, a.eta, ku, kv);
    s_rhs0[tid][0] = ku;
    s_rhs0[tid][1] = kv;
  }
  if (tid < 12 * TYB) {
    const int row = tid / 12, rem = tid % 12, gf = rem / 6, pt = rem % 6;
    if (i1s + row < g.N1 + NG) {
      const T *box = gf == 0 ? bu : bv;
      const int c0 = NG + 1 + pt, by = row + NG, sh = pt < 1 ? 1 : pt;  // stencil shift -max(1,pt)
      T d0 = 0, d1 = 0;
      if (a.bc_kind == NRPYELL_OUTER_BC_RADIATION) {
#pragma unroll
        for (int j = 0; j < 11; j++) {
          d0 = fma(s_fd[sh * 11 + j], box[by * NBX + c0 + (j - 5 - sh)], d0);
          d1 = fma((T)c_fd1_shift[0][j], box[(by + j - 5) * NBX + c0], d1);
        }
      } else {
        // the Einstein Toolkit thorn's 7-point stencils (conformally_flat_BBH_apply_bcs_sommerfeld.h:
        // 49-105): backward one-sided along x0 (columns c0-6..c0, c0-6 = pt >= 0); along x1
        // forward one-sided on the first interior row (rows by..by+6 < BYB), centred elsewhere
        const T *col = box + by * NBX + c0;
        d0 = (T)(-49.0 / 20.0) * col[0];
#pragma unroll
        for (int k = 1; k < 7; k++) d0 = fma((T)c_fd1_onesided7[k], col[-k], d0);
        d0 = -d0;
        if (i1s + row == NG) {
          d1 = (T)(-49.0 / 20.0) * col[0];
#pragma unroll
          for (int k = 1; k < 7; k++) d1 = fma((T)c_fd1_onesided7[k], col[k * NBX], d1);
        } else {
          d1 = (T)(1.0 / 60.0) * col[3 * NBX];
          d1 = fma((T)(-3.0 / 20.0), col[2 * NBX], d1);
          d1 = fma((T)(3.0 / 4.0), col[NBX], d1);
          d1 = fma((T)(-3.0 / 4.0), col[-NBX], d1);
          d1 = fma((T)(3.0 / 20.0), col[-2 * NBX], d1);
          d1 = fma((T)(-1.0 / 60.0), col[-3 * NBX], d1);
        }
      }
      d0 *= (T)g.invdx0;
      d1 *= (T)g.invdx1;
      const T dr = s_geo[pt][row][1] * d0 + s_geo[pt][row][2] * d1;
      const T cc = gf == 0 ? a.cOB_u : a.cOB_v;
      s_A[row][gf][pt] = -cc * (dr + box[by * NBX + c0] * s_geo[pt][row][0]);
    }
  }
  __syncthreads();
  if (od_on) {
    const T rhs = fma(s_geo[od_pt][od_row][3], s_rhs0[od_row][od_gf] - s_A[od_row][od_gf][0],
                      s_A[od_row][od_gf][od_pt]);
    T o;
    T rt_ = od_rt;
    stage_update<T, STAGE>(rhs, a.dt, od_yn, rt_, o);
    if (STAGE != 4) st32(a.rt, od_p, rt_);
    st32(a.out, od_p, o);
    for_each_image_in_planes(g, ilast + od_pt, i1s + od_row, i2,
                             [&](long d) { st32(a.out, od_gf * nt + (unsigned)d, o); });
  }
}

// One configuration: the nbc_rows*N2 BC blocks first (dispatched first, spread over the XCDs),
// then the ntile0*ntile1*(N2/PL) interior blocks in XCD-aware order.
template <typename T, int STAGE, int TX, int PL>
__global__ void __launch_bounds__(NRPYELL_BLOCK) substep_kernel(SubstepArgs<T> a) {
  const unsigned L = blockIdx.x, nb = (unsigned)(a.nbc_rows * a.g.N2);
  if (L < nb) {
    const unsigned pln = div_by(L, a.nbc_mul);
    bc_body<T, STAGE>(a, (int)(L - pln * (unsigned)a.nbc_rows), (int)pln);
    return;
  }
  const unsigned ni = gridDim.x - nb, q = xcd_logical(L - nb, ni), ng = (unsigned)(a.g.N2 / PL);
  const unsigned tile = div_by(q, a.ng_mul);
  interior_body<T, STAGE, TX, PL>(a, (int)tile, (int)(q - tile * ng), q, ni);
}

// Many configurations (uneven grids) in one launch: logical block q works for configuration
// block_cfg[q], as that configuration's block q - block_first[cfg] (its BC blocks first, then
// its interior blocks). Configurations that have converged return at once (their own done
// flag) and are dropped from the launch when the host compacts the batch.
template <typename T, int STAGE, int TX, int PL>
__global__ void __launch_bounds__(NRPYELL_BLOCK) substep_batched_kernel(const SubstepArgs<T> *__restrict__ cfgs,
                                                              const int *__restrict__ block_cfg,
                                                              const int *__restrict__ block_first) {
  const unsigned q = xcd_logical(blockIdx.x, gridDim.x);
  const int c = block_cfg[q];
  const SubstepArgs<T> a = cfgs[c];
  if (*a.done) return;  // leave before issuing any load
  const unsigned local = q - block_first[c], nb = (unsigned)(a.nbc_rows * a.g.N2);
  if (local < nb) {
    const unsigned pln = div_by(local, a.nbc_mul);
    bc_body<T, STAGE>(a, (int)(local - pln * (unsigned)a.nbc_rows), (int)pln);
    return;
  }
  const unsigned li = local - nb, ng = (unsigned)(a.g.N2 / PL);
  const unsigned tile = div_by(li, a.ng_mul);
  interior_body<T, STAGE, TX, PL>(a, (int)tile, (int)(li - tile * ng), li, a.ntile0 * a.ntile1 * ng);
}

template <typename T, int STAGE>
static void launch_stage(const SubstepArgs<T> &a, hipStream_t st) {
  dim3 grd(a.nbc_rows * a.g.N2 + a.ntile0 * a.ntile1 * (a.g.N2 / a.pl)), blk(NRPYELL_BLOCK);
  if (a.pl == 2) {
    switch (a.tx) {
      case 64: hipLaunchKernelGGL((substep_kernel<T, STAGE, 64, 2>), grd, blk, 0, st, a); break;
      default: hipLaunchKernelGGL((substep_kernel<T, STAGE, 32, 2>), grd, blk, 0, st, a); break;
    }
    return;
  }
  switch (a.tx) {
    case 16: hipLaunchKernelGGL((substep_kernel<T, STAGE, 16, 1>), grd, blk, 0, st, a); break;
    case 64: hipLaunchKernelGGL((substep_kernel<T, STAGE, 64, 1>), grd, blk, 0, st, a); break;
    default: hipLaunchKernelGGL((substep_kernel<T, STAGE, 32, 1>), grd, blk, 0, st, a); break;
  }
}

template <typename T, int STAGE>
static void launch_stage_batched(const SubstepArgs<T> *cfgs, const int *block_cfg, const int *block_first,
                                 int nblocks, int tx, int pl, hipStream_t st) {
  dim3 grd(nblocks), blk(NRPYELL_BLOCK);
  if (pl == 2) {
    switch (tx) {
      case 64: hipLaunchKernelGGL((substep_batched_kernel<T, STAGE, 64, 2>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
      default: hipLaunchKernelGGL((substep_batched_kernel<T, STAGE, 32, 2>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    }
    return;
  }
  switch (tx) {
    case 16: hipLaunchKernelGGL((substep_batched_kernel<T, STAGE, 16, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    case 64: hipLaunchKernelGGL((substep_batched_kernel<T, STAGE, 64, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    default: hipLaunchKernelGGL((substep_batched_kernel<T, STAGE, 32, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
  }
}

template <typename T>
void launch_substep_batched(const SubstepArgs<T> *cfgs, const int *block_cfg, const int *block_first, int nblocks,
                            int tx, int pl, int stage, hipStream_t st) {
  switch (stage) {
    case 1: launch_stage_batched<T, 1>(cfgs, block_cfg, block_first, nblocks, tx, pl, st); break;
    case 2: launch_stage_batched<T, 2>(cfgs, block_cfg, block_first, nblocks, tx, pl, st); break;
    case 3: launch_stage_batched<T, 3>(cfgs, block_cfg, block_first, nblocks, tx, pl, st); break;
    default: launch_stage_batched<T, 4>(cfgs, block_cfg, block_first, nblocks, tx, pl, st); break;
  }
}
template void launch_substep_batched<double>(const SubstepArgs<double> *, const int *, const int *, int, int, int,
                                             int, hipStream_t);
template void launch_substep_batched<float>(const SubstepArgs<float> *, const int *, const int *, int, int, int,
                                            int, hipStream_t);

template <typename T>
void launch_substep(const SubstepArgs<T> &a, int stage, hipStream_t st) {
  switch (stage) {
    case 1: launch_stage<T, 1>(a, st); break;
    case 2: launch_stage<T, 2>(a, st); break;
    case 3: launch_stage<T, 3>(a, st); break;
    default: launch_stage<T, 4>(a, st); break;
  }
}
template void launch_substep<double>(const SubstepArgs<double> &, int, hipStream_t);
template void launch_substep<float>(const SubstepArgs<float> &, int, hipStream_t);
