// nrpyell_substep2.hip -- two RK4 substeps per launch (stages 1+2, then stages 3+4): the path of
// the resident solver with the radiation outer BC, in FP64 and in the FP32 bulk phase of the
// mixed-precision solve (templated on the arithmetic type T).
//
// RK4 on  u_t = v - eta u,  v_t = c^2 (lap u + AA^2 / (8 (psi+u)^7))  couples neighbouring points
// only through the stencil of u: the next stage's u is pointwise,
//   u_{s+1} = y_u + a_s dt (v_s - eta u_s),
// so a block that reads stage s's (u, v) on its tile plus the 5-point halo can form u_{s+1} on the
// whole halo itself and evaluate stage s+1's stencil without waiting for its neighbours. One
// launch then does two substeps with no redundant stencil work (only the pointwise u_{s+1} on the
// halo is recomputed), and an RK4 step is two dependent launches instead of four. Per point the
// arithmetic is that of the one-substep kernel (same stage updates, same stencil sums, same
// radiation terms): the two paths give bit-identical states.
//
// Interior blocks: tiles of TX (i0) x TY (i1) points on PL i2 planes, aligned so that the last
// tile ends at the last interior i0 (ilast). Only that tile's stencils reach the outer ghost
// layers; it forms u_{s+1} there from stage s's radiation condition (the same closed form as the
// BC blocks, from its LDS box).
//
// BC blocks (rows r0..r0+TYB-1 of one plane) produce the outer-layer values of stage s+1. That
// needs stage s+1 on the outer columns with i1 neighbours +-5 rows, hence stage s on +-5 more rows
// and stage s's inputs on +-10 rows: rows outside [NG, N1+NG) are taken through the SinhSymTP
// reflection (row -> mirrored row, phi -> phi + pi) as "virtual rows", every quantity on them
// evaluated at the source point in the source's own orientation (the i1 arm read with the
// reflection's direction), so it equals the source's own value bit for bit.
//
// Buffers of one step n -> n+1 (the solver alternates y between two arrays A_j, A_(1-j)):
//   launch 1 (stages 1+2): in = y = A_j,  writes k_odd = y3 = y + dt/2 k2 and rt = dt/6 k1 + dt/3 k2
//   launch 2 (stages 3+4): in = k_odd, y = A_j, rt; writes A_(1-j) = y + (rt + dt/3 k3) + dt/6 k4
// (Storing stage 4's u from launch 1, so that launch 2 reads two arrays instead of three on its
// boxes and i2 arms, measured no faster: launch 2 gained what launch 1 lost.)
// y_n is never overwritten within its step, so when stage 1's fused residual check (reduced by
// the second launch's block 0) ends the solve, A_j still holds the state it checked.
#include <hip/hip_runtime.h>
#include <math.h>

#include "nrpyell_common.h"
#include "nrpyell_internal.h"
#include "nrpyell_stencil.h"

#define NG NRPYELL_NG


// Stage s's contribution to the next stage's input, y + a_s dt k_s, with the FMA placement of
// stage_update (MoL_step_forward_in_time.c:44-45, :88-89, :129-130).
template <typename T, int STAGE>
__device__ __forceinline__ T next_input(T k, T dt, T yn) {
  if (STAGE == 3) return fma(dt, k, yn);
  return fma((T)0.5, dt * k, yn);
}

// Radiation term A = -c (d_r f + f / r) at outer layer pt (pt = 0: the last interior point) of
// apply_bcs_outerradiation_and_inner.c:290-408, as bc_body evaluates it: f(di0, di1) is f at
// offsets (di0, di1) from the point, d_0 the 11-point stencil shifted by -max(1, pt) toward the
// interior, d_1 the centred one; geo = (1/r, dxx0/dr, dxx1/dr, (r_0/r)^3). fd is the block's LDS
// copy of c_fd1_shift (load_fd_table): pt differs between lanes, and a lane-indexed read of the
// constant table is a global-memory round trip in the middle of the block's dependency chain.
template <typename T, class F>
__device__ __forceinline__ T radiation_A(F &&f, int pt, const T geo[4], T c, T invdx0, T invdx1, const T *fd) {
  const int sh = pt < 1 ? 1 : pt;
  T d0 = 0, d1 = 0;
#pragma unroll
  for (int j = 0; j < 11; j++) {
    d0 = fma(fd[sh * 11 + j], f(j - 5 - sh, 0), d0);
    d1 = fma((T)c_fd1_shift[0][j], f(0, j - 5), d1);
  }
  d0 *= invdx0;
  d1 *= invdx1;
  const T dr = geo[1] * d0 + geo[2] * d1;
  return -c * (dr + f(0, 0) * geo[0]);
}

// Source row, source plane and i1 orientation of a row index that may lie beyond the grid's
// stored ghost rows: the SinhSymTP reflection xx1 -> -xx1 (or 2 pi - xx1) with a half turn in phi.
// Rows more than N1 - 10 beyond the last interior row occur only in the unused tail of a BC
// block on a grid with N1 < NRPYELL_BC_ROWS + 10; they are clamped into the grid.
__device__ __forceinline__ void vrow_source(const GridDims &g, int vr, int i2, int &row, int &pl, int &dir) {
  if (vr < NG) {
    row = min(2 * NG - 1 - vr, g.N1 + NG - 1);
    pl = wrap2_near(g, i2 + g.N2 / 2);
    dir = -1;
  } else if (vr >= g.N1 + NG) {
    row = max(2 * (g.N1 + NG) - 1 - vr, NG);
    pl = wrap2_near(g, i2 + g.N2 / 2);
    dir = -1;
  } else {
    row = vr;
    pl = i2;
    dir = 1;
  }
}

// ---------------------------------------------------------------------------------------
// Interior blocks.
// ---------------------------------------------------------------------------------------
// LDS (elements of T) of an interior block and of a BC block; the kernel holds one buffer of the
// larger size and each body lays its arrays out in it (the block sums have their own array).
template <int TX, int PL>
constexpr int interior2_lds() {
  constexpr int NT = NRPYELL_BLOCK, TY = NT / TX, BOX = (TX + 2 * NG) * (TY + 2 * NG), NR = (BOX + NT - 1) / NT;
  return 2 * PL * NR * NT + 24 * TY + NRPYELL_FD_TABLE;
}
constexpr int bc2_lds() {
  constexpr int NT = NRPYELL_BLOCK, TYB = NRPYELL_PAIR_BC_ROWS, NR1 = TYB + 2 * NG;
  constexpr int NRB = ((TYB + 4 * NG) * 17 + NT - 1) / NT;
  return 2 * NRB * NT + NR1 * (24 + 10 + 24 + 2 + 12) + TYB * (10 + 30 + 8 + 10 + 2 + 12) + NRPYELL_FD_TABLE;
}

template <typename T, int PAIR, int TX, int PL>
__device__ __forceinline__ void interior2_body(const SubstepArgs<T> &a, const int tile, const int pgroup,
                                               const unsigned bid, const unsigned nblk, T *lds, double *s_red) {
  constexpr int S0 = PAIR;
  constexpr int NT = NRPYELL_BLOCK, TY = NT / TX;
  constexpr int BX = TX + 2 * NG, BY = TY + 2 * NG, BOX = BX * BY, NR = (BOX + NT - 1) / NT;
  constexpr int NW = PL + 2 * NG;
  constexpr int NBI = 5 * TY * PL;  // outer points of the last tile's rows
  constexpr int NGEO = (24 * TY + NT - 1) / NT;
  static_assert(NBI <= NT, "one outer point per thread");
  // stage-s u and stage-(s+1) u on the box of each plane, the last tile's geometry, block sums
  T(*sx)[NR * NT] = reinterpret_cast<T(*)[NR * NT]>(lds);
  T(*s1)[NR * NT] = reinterpret_cast<T(*)[NR * NT]>(lds + PL * NR * NT);
  T(*s_geo)[TY][4] = reinterpret_cast<T(*)[TY][4]>(lds + 2 * PL * NR * NT);
  T *s_fd = lds + 2 * PL * NR * NT + 24 * TY;

  const GridDims &g = a.g;
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int t1 = (int)div_by((unsigned)tile, a.nt0_mul), t0 = tile - t1 * a.ntile0;
  const int i2a = NG + PL * pgroup;
  const int i0s = a.i0_first + t0 * TX, i1s = NG + t1 * TY;
  const int i0 = i0s + tx, i1 = i1s + ty;
  const int ilast = g.N0 + NG - 1;
  const bool valid = i0 >= NG && i0 <= ilast && i1 < g.N1 + NG;
  const bool last_tile = t0 == a.ntile0 - 1;  // block-uniform
  const unsigned nt = (unsigned)g.ntot, s1s = g.G0, s2 = (unsigned)(g.G0 * g.G1);
  const bool fuse_l2 = (PAIR == 1) && (a.l2_partial != nullptr);
  const T eta = a.eta, dt = a.dt;
  const int done = *a.done;

  // ---- loads: unconditional, from clamped indices (values a stencil never reaches are loaded
  // and not read). For stages 1+2 the stage input is y itself.
  const int i0c = min(max(i0, NG), ilast), i1c = min(i1, g.N1 + NG - 1);
  const unsigned k2 = (unsigned)i1c * s1s + i0c;
  T bxu[PL][NR], bxv[PL][NR], byu[PL][NR];
  T wxu[NW], wxv[NW], wyu[NW];
  T yv[PL], psi[PL], AAc[PL], rtu[PL], rtv[PL], K[5];
  double w2 = 0.0;
#pragma unroll
  for (int q = 0; q < PL; q++) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int k = tid + NT * r;
      const int bx = k % BX, by = k / BX;
      const int j0 = min(max(i0s - NG + bx, 0), g.G0 - 1), j1 = min(i1s - NG + by, g.G1 - 1);
      const unsigned p = (unsigned)(i2a + q) * s2 + (unsigned)j1 * s1s + j0;
      bxu[q][r] = ld32(a.in, p);
      bxv[q][r] = ld32(a.in, nt + p);
      byu[q][r] = PAIR == 1 ? bxu[q][r] : ld32(a.yn, p);
    }
    if (q == 0) {
#pragma unroll
      for (int jj = 0; jj < NW; jj++) {
        const unsigned pw = (unsigned)wrap2_near(g, i2a - NG + jj) * s2 + k2;
        wxu[jj] = ld32(a.in, pw);
        wxv[jj] = ld32(a.in, nt + pw);
        wyu[jj] = PAIR == 1 ? wxu[jj] : ld32(a.yn, pw);
      }
      T kw;
      ld2x32(a.coef, 6 * k2, K[0], K[1]);
      ld2x32(a.coef, 6 * k2 + 2, K[2], K[3]);
      ld2x32(a.coef, 6 * k2 + 4, K[4], kw);
      if (fuse_l2) w2 = (double)kw;
    }
    const unsigned p = (unsigned)(i2a + q) * s2 + k2;
    yv[q] = PAIR == 1 ? wxv[NG + q] : ld32(a.yn, nt + p);
    ld2x32(a.src, 2 * p, psi[q], AAc[q]);
    rtu[q] = 0;
    rtv[q] = 0;
    if (PAIR == 3) {
      rtu[q] = ld32((const T *)a.rt, p);
      rtv[q] = ld32((const T *)a.rt, nt + p);
    }
  }
  // the last tile's outer points: plane bq, row brow, layer bpt; y_u there and v at ilast
  const int bq = min(tid / (5 * TY), PL - 1), brow = (tid / 5) % TY, bpt = tid % 5 + 1;
  const bool b_on = tid < NBI && i1s + brow < g.N1 + NG;
  T b_yu = 0, b_xv0 = 0;
  T geo_r[NGEO], fdv = 0;
  if (last_tile) {
    fdv = (T)load_fd_table(tid);
    const unsigned bp = (unsigned)(i2a + bq) * s2 + (unsigned)min(i1s + brow, g.N1 + NG - 1) * s1s;
    b_yu = ld32(a.yn, bp + ilast + bpt);
    b_xv0 = ld32(a.in, nt + bp + ilast);
#pragma unroll
    for (int r = 0; r < NGEO; r++) {
      const int k = min(tid + NT * r, 24 * TY - 1);
      const int pt = k / (4 * TY), row = (k / 4) % TY, comp = k % 4;
      geo_r[r] = ld32(a.geo, ((unsigned)pt * g.G1 + min(i1s + row, g.G1 - 1)) * 4 + comp);
    }
  }
  if (done) return;  // converged: uniform across the grid

  // ---- stage s's u and the pointwise stage-(s+1) u on the whole box
#pragma unroll
  for (int q = 0; q < PL; q++) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      sx[q][tid + NT * r] = bxu[q][r];
      s1[q][tid + NT * r] = next_input<T, S0>(fma(-eta, bxu[q][r], bxv[q][r]), dt, byu[q][r]);
    }
  }
  if (last_tile) {
#pragma unroll
    for (int r = 0; r < NGEO; r++)
      if (tid + NT * r < 24 * TY) (&s_geo[0][0][0])[tid + NT * r] = geo_r[r];
    if (tid < NRPYELL_FD_TABLE) s_fd[tid] = fdv;
  }
  __syncthreads();

  // ---- the last tile: stage-(s+1) u on the outer layers of its rows, from stage s's radiation
  // condition rhs_pt = A_pt + (r_0/r_pt)^3 (rhs_0 - A_0), rhs_0 = v - eta u at ilast
  if (last_tile) {
    if (b_on) {
      const int cl = ilast - i0s + NG, by = brow + NG;  // box column of ilast, box row
      const T *box = sx[bq];
      const T Ap = radiation_A<T>([&](int d0, int d1) { return box[(by + d1) * BX + cl + bpt + d0]; }, bpt,
                                  s_geo[bpt][brow], a.cOB_u, (T)g.invdx0, (T)g.invdx1, s_fd);
      const T A0 = radiation_A<T>([&](int d0, int d1) { return box[(by + d1) * BX + cl + d0]; }, 0,
                                  s_geo[0][brow], a.cOB_u, (T)g.invdx0, (T)g.invdx1, s_fd);
      const T rhs0 = fma(-eta, box[by * BX + cl], b_xv0);
      const T rhs = fma(s_geo[bpt][brow][3], rhs0 - A0, Ap);
      s1[bq][by * BX + cl + bpt] = next_input<T, S0>(rhs, dt, b_yu);
    }
    __syncthreads();
  }

  // ---- both stages at the tile points
  double l2acc = 0.0;
  const int cb = (ty + NG) * BX + tx + NG;
#pragma unroll
  for (int q = 0; q < PL; q++) {
    if (valid) {
      const int i2 = i2a + q;
      const unsigned p = (unsigned)i2 * s2 + k2;
      T xm[6], xp[6], ym[6], yp[6], zm[6], zp[6];
      // stage s
      const T uu = wxu[NG + q], vv = wxv[NG + q];
#pragma unroll
      for (int k = 1; k <= 5; k++) {
        xm[k] = sx[q][cb - k]; xp[k] = sx[q][cb + k];
        ym[k] = sx[q][cb - k * BX]; yp[k] = sx[q][cb + k * BX];
        zm[k] = wxu[NG + q - k]; zp[k] = wxu[NG + q + k];
      }
      T ku0, kv0;
      point_rhs<T>(uu, vv, psi[q], AAc[q], K, xm, xp, ym, yp, zm, zp, eta, ku0, kv0);
      if (fuse_l2) l2acc += w2 * (double)kv0 * (double)kv0;  // w / c^4 x (c^2 R)^2
      // stage s+1
      const T u1 = s1[q][cb];
      const T v1 = next_input<T, S0>(kv0, dt, yv[q]);
#pragma unroll
      for (int k = 1; k <= 5; k++) {
        xm[k] = s1[q][cb - k]; xp[k] = s1[q][cb + k];
        ym[k] = s1[q][cb - k * BX]; yp[k] = s1[q][cb + k * BX];
        zm[k] = next_input<T, S0>(fma(-eta, wxu[NG + q - k], wxv[NG + q - k]), dt, wyu[NG + q - k]);
        zp[k] = next_input<T, S0>(fma(-eta, wxu[NG + q + k], wxv[NG + q + k]), dt, wyu[NG + q + k]);
      }
      T ku1, kv1;
      point_rhs<T>(u1, v1, psi[q], AAc[q], K, xm, xp, ym, yp, zm, zp, eta, ku1, kv1);
      T ou, ov;
      if (PAIR == 1) {
        T ru = (T)(1.0 / 6.0) * (dt * ku0), rv = (T)(1.0 / 6.0) * (dt * kv0);
        ru = fma((T)(1.0 / 3.0), dt * ku1, ru);
        rv = fma((T)(1.0 / 3.0), dt * kv1, rv);
        ou = fma((T)0.5, dt * ku1, wyu[NG + q]);
        ov = fma((T)0.5, dt * kv1, yv[q]);
        st32(a.rt, p, ru);
        st32(a.rt, nt + p, rv);
      } else {
        const T ru = fma((T)(1.0 / 3.0), dt * ku0, rtu[q]), rv = fma((T)(1.0 / 3.0), dt * kv0, rtv[q]);
        ou = wyu[NG + q] + fma((T)(1.0 / 6.0), dt * ku1, ru);
        ov = yv[q] + fma((T)(1.0 / 6.0), dt * kv1, rv);
      }
      st32(a.out, p, ou);
      st32(a.out, nt + p, ov);
      for_each_image_in_planes(g, i0, i1, i2, [&](long d) {
        st32(a.out, (unsigned)d, ou);
        st32(a.out, nt + (unsigned)d, ov);
      });
    }
  }

  // ---- residual L2 of y_n (stage 1's input): one partial per block; block 0 of the stages-3+4
  // launch that follows reduces them in block order and updates the convergence state
  if (fuse_l2) {
    for (int o = 32; o > 0; o >>= 1) l2acc += __shfl_down(l2acc, o, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = l2acc;
    __syncthreads();
    if (tid == 0) a.l2_partial[bid] = wave_sum<NT / 64>(s_red);
  }
  if (PAIR == 3 && a.l2_partial != nullptr && bid == 0) {
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
      record_check(st, l, a.l2_target);
    }
  }
}

// ---------------------------------------------------------------------------------------
// BC blocks: the outer layers of rows r0 .. r0+TYB-1 of one plane after stage s+1.
//   box:  stage-s (u, v) on rows r0-10 .. r0+TYB+9 (virtual rows) x columns ilast-11 .. ilast+5
//   R1:   rows r0-5 .. r0+TYB+4 x columns ilast-6 .. ilast+5: stage-(s+1) (u, v), from stage s's
//         stencil (interior columns) and radiation condition (outer columns)
//   then stage s+1's RHS at ilast and radiation terms on the tile rows, and the outputs.
// Every global load is issued before the first barrier (one memory round trip per block); the
// phases after it work from registers and LDS.
// ---------------------------------------------------------------------------------------
// Stage-s stencil point t of a BC block: (R1 row e, column col); t < 6 TYB: a tile row (e = NG
// + t / 6) and a column ilast-6 .. ilast-1; then column ilast on R1 rows 0 .. TYB+2NG-1.
__device__ __forceinline__ void r1_stencil_point(int t, int ilast, int &e, int &col) {
  constexpr int TYB = NRPYELL_PAIR_BC_ROWS;
  if (t < 6 * TYB) {
    e = NG + t / 6;
    col = ilast - 6 + t % 6;
  } else {
    e = t - 6 * TYB;
    col = ilast;
  }
}

template <typename T, int PAIR>
__device__ __forceinline__ void bc2_body(const SubstepArgs<T> &a, const int rb, const int plane, T *lds) {
  constexpr int S0 = PAIR;
  constexpr int NT = NRPYELL_BLOCK, TYB = NRPYELL_PAIR_BC_ROWS;
  constexpr int NVR = TYB + 4 * NG, NC = 17, NBOX = NVR * NC, NRB = (NBOX + NT - 1) / NT;
  // stage-s stencils: columns ilast-6 .. ilast-1 on the tile rows (the one-sided d_0 of stage
  // s+1's radiation terms reads them there only) and column ilast on every R1 row (d_1 there)
  constexpr int NR1 = TYB + 2 * NG, NI = 6 * TYB + NR1, NIR = (NI + NT - 1) / NT;
  constexpr int NA0 = NR1 * 12, NAR = (NA0 + NT - 1) / NT;  // stage-s radiation terms
  constexpr int NO0 = NR1 * 10, NOR = (NO0 + NT - 1) / NT;  // stage-s outer values
  constexpr int NGE = NR1 * 24, NGER = (NGE + NT - 1) / NT;  // geometry of the R1 rows
  constexpr int NARM = TYB * 30, NARMR = (NARM + NT - 1) / NT;  // i2 arms at (ilast, tile rows)
  static_assert(TYB * 12 <= NT && TYB * 10 <= NT && TYB * 8 <= NT, "one tile-row item per thread");
  T *bxu = lds, *bxv = lds + NRB * NT;  // [vr - (r0-10)][c - (ilast-11)]
  T *q = lds + 2 * NRB * NT;
  auto take = [&](int n) { T *r = q; q += n; return r; };
  T(*sgeo)[6][4] = reinterpret_cast<T(*)[6][4]>(take(NR1 * 24));  // at each R1 row's source row
  T(*syo)[2][5] = reinterpret_cast<T(*)[2][5]>(take(NR1 * 10));   // y at the outer points of R1
  T(*srti)[2][5] = reinterpret_cast<T(*)[2][5]>(take(TYB * 10));  // (stages 3+4) running total there
  T(*sarm)[10][3] = reinterpret_cast<T(*)[10][3]>(take(TYB * 30));  // stage-s u, v, y_u on i2 arms at ilast
  T(*skl)[8] = reinterpret_cast<T(*)[8]>(take(TYB * 8));          // coefficients, source at (ilast, row)
  T(*ru1)[12] = reinterpret_cast<T(*)[12]>(take(NR1 * 12));       // stage-(s+1) u on R1
  T(*rv1)[12] = reinterpret_cast<T(*)[12]>(take(NR1 * 12));       // stage-(s+1) v on R1
  T(*sk0)[2] = reinterpret_cast<T(*)[2]>(take(NR1 * 2));          // stage-s k at (ilast, R1 row)
  T(*sA0)[2][6] = reinterpret_cast<T(*)[2][6]>(take(NR1 * 12));   // stage-s radiation terms, R1 rows
  T(*srto)[2][5] = reinterpret_cast<T(*)[2][5]>(take(TYB * 10));  // running total at the outputs
  T(*sk1)[2] = reinterpret_cast<T(*)[2]>(take(TYB * 2));
  T(*sA1)[2][6] = reinterpret_cast<T(*)[2][6]>(take(TYB * 12));
  T *s_fd = take(NRPYELL_FD_TABLE);

  const GridDims &g = a.g;
  const int tid = threadIdx.x;
  const unsigned nt = (unsigned)g.ntot, s1s = g.G0, s2 = (unsigned)(g.G0 * g.G1);
  const int ilast = g.N0 + NG - 1, c_lo = ilast - 11;
  const int r0 = NG + rb * TYB, i2 = NG + plane;
  const T eta = a.eta, dt = a.dt, ix0 = (T)g.invdx0, ix1 = (T)g.invdx1;
  const int done = *a.done;

  // ---- loads
  T lxu[NRB], lxv[NRB];
#pragma unroll
  for (int r = 0; r < NRB; r++) {
    const int k = min(tid + NT * r, NBOX - 1), e = k / NC, c = k % NC;
    int row, pl, dir;
    vrow_source(g, r0 - 2 * NG + e, i2, row, pl, dir);
    const unsigned p = (unsigned)pl * s2 + (unsigned)row * s1s + (c_lo + c);
    lxu[r] = ld32(a.in, p);
    lxv[r] = ld32(a.in, nt + p);
  }
  // this thread's stage-s stencil points on R1 (columns ilast-6 .. ilast)
  T iK[NIR][5], ipsi[NIR], iAAc[NIR], izm[NIR][6], izp[NIR][6], iyu[NIR], iyv[NIR];
#pragma unroll
  for (int it = 0; it < NIR; it++) {
    int e, col;
    r1_stencil_point(min(tid + NT * it, NI - 1), ilast, e, col);
    int row, pl, dir;
    vrow_source(g, r0 - NG + e, i2, row, pl, dir);
    const unsigned k2 = (unsigned)row * s1s + col, p = (unsigned)pl * s2 + k2;
    T kw;
    ld2x32(a.coef, 6 * k2, iK[it][0], iK[it][1]);
    ld2x32(a.coef, 6 * k2 + 2, iK[it][2], iK[it][3]);
    ld2x32(a.coef, 6 * k2 + 4, iK[it][4], kw);
    ld2x32(a.src, 2 * p, ipsi[it], iAAc[it]);
#pragma unroll
    for (int k = 1; k <= 5; k++) {
      izm[it][k] = ld32(a.in, (unsigned)wrap2_near(g, pl - k) * s2 + k2);
      izp[it][k] = ld32(a.in, (unsigned)wrap2_near(g, pl + k) * s2 + k2);
    }
    iyu[it] = PAIR == 1 ? (T)0 : ld32(a.yn, p);
    iyv[it] = PAIR == 1 ? (T)0 : ld32(a.yn, nt + p);
  }
  T lgeo[NGER];
#pragma unroll
  for (int r = 0; r < NGER; r++) {
    const int k = min(tid + NT * r, NGE - 1), e = k / 24, pt = (k / 4) % 6, c = k % 4;
    int row, pl, dir;
    vrow_source(g, r0 - NG + e, i2, row, pl, dir);
    lgeo[r] = ld32(a.geo, ((unsigned)pt * g.G1 + row) * 4 + c);
  }
  T lyo[NOR], lrt = 0;
#pragma unroll
  for (int r = 0; r < NOR; r++) {
    const int k = min(tid + NT * r, NO0 - 1), e = k / 10, gf = (k % 10) / 5, pt = k % 5 + 1;
    int row, pl, dir;
    vrow_source(g, r0 - NG + e, i2, row, pl, dir);
    lyo[r] = ld32(a.yn, (unsigned)pl * s2 + (unsigned)row * s1s + (ilast + pt) + gf * nt);
  }
  if (PAIR == 3) {
    const int k = min(tid, TYB * 10 - 1), tr = k / 10, gf = (k % 10) / 5, pt = k % 5 + 1;
    lrt = ld32((const T *)a.rt, i2 * s2 + (unsigned)min(r0 + tr, g.N1 + NG - 1) * s1s + (ilast + pt) + gf * nt);
  }
  T larm[NARMR];
#pragma unroll
  for (int r = 0; r < NARMR; r++) {
    const int k = min(tid + NT * r, NARM - 1), tr = k / 30, j = (k / 3) % 10, f = k % 3;
    const int kk = j < 5 ? -(j + 1) : j - 4;  // plane offsets -1..-5, +1..+5
    const unsigned pa = (unsigned)wrap2_near(g, i2 + kk) * s2 + (unsigned)min(r0 + tr, g.N1 + NG - 1) * s1s + ilast;
    larm[r] = f == 0 ? ld32(a.in, pa) : f == 1 ? ld32(a.in, nt + pa) : ld32(a.yn, pa);
  }
  const T lfd = (T)load_fd_table(tid);
  T lkl = 0;
  if (tid < TYB * 8) {
    const int tr = tid / 8, c = tid % 8;
    const unsigned k2 = (unsigned)min(r0 + tr, g.N1 + NG - 1) * s1s + ilast;
    lkl = c < 6 ? ld32(a.coef, 6 * k2 + c) : ld32(a.src, 2 * (i2 * s2 + k2) + (c - 6));
  }
  if (done) return;

  // ---- everything to LDS
#pragma unroll
  for (int r = 0; r < NRB; r++) {
    bxu[tid + NT * r] = lxu[r];
    bxv[tid + NT * r] = lxv[r];
  }
#pragma unroll
  for (int r = 0; r < NGER; r++)
    if (tid + NT * r < NGE) (&sgeo[0][0][0])[tid + NT * r] = lgeo[r];
#pragma unroll
  for (int r = 0; r < NOR; r++)
    if (tid + NT * r < NO0) (&syo[0][0][0])[tid + NT * r] = lyo[r];
  if (PAIR == 3 && tid < TYB * 10) (&srti[0][0][0])[tid] = lrt;
#pragma unroll
  for (int r = 0; r < NARMR; r++)
    if (tid + NT * r < NARM) (&sarm[0][0][0])[tid + NT * r] = larm[r];
  if (tid < TYB * 8) (&skl[0][0])[tid] = lkl;
  if (tid < NRPYELL_FD_TABLE) s_fd[tid] = lfd;
  __syncthreads();
  // box value at virtual row vr (absolute) and column c (absolute)
  auto BX_U = [&](int vr, int c) { return bxu[(vr - (r0 - 2 * NG)) * NC + (c - c_lo)]; };
  auto BX_V = [&](int vr, int c) { return bxv[(vr - (r0 - 2 * NG)) * NC + (c - c_lo)]; };

  // ---- stage s at the R1 points of columns ilast-6 .. ilast
#pragma unroll
  for (int it = 0; it < NIR; it++) {
    const int t = tid + NT * it;
    if (t < NI) {
      int e, col;
      r1_stencil_point(t, ilast, e, col);
      const int vr = r0 - NG + e;
      int row, pl, dir;
      vrow_source(g, vr, i2, row, pl, dir);
      T xm[6], xp[6], ym[6], yp[6];
      const T uu = BX_U(vr, col), vv = BX_V(vr, col);
#pragma unroll
      for (int k = 1; k <= 5; k++) {
        xm[k] = BX_U(vr, col - k); xp[k] = BX_U(vr, col + k);
        ym[k] = BX_U(vr - dir * k, col); yp[k] = BX_U(vr + dir * k, col);  // the source's orientation
      }
      T ku, kv;
      point_rhs<T>(uu, vv, ipsi[it], iAAc[it], iK[it], xm, xp, ym, yp, izm[it], izp[it], eta, ku, kv);
      ru1[e][col - (ilast - 6)] = next_input<T, S0>(ku, dt, PAIR == 1 ? uu : iyu[it]);
      rv1[e][col - (ilast - 6)] = next_input<T, S0>(kv, dt, PAIR == 1 ? vv : iyv[it]);
      if (col == ilast) {
        sk0[e][0] = ku;
        sk0[e][1] = kv;
      }
    }
  }
  // ---- stage s's radiation terms A_pt (pt 0..5, both gridfunctions) on the R1 rows
#pragma unroll
  for (int it = 0; it < NAR; it++) {
    const int t = tid + NT * it;
    if (t < NA0) {
      const int e = t / 12, gf = (t % 12) / 6, pt = t % 6, vr = r0 - NG + e;
      int row, pl, dir;
      vrow_source(g, vr, i2, row, pl, dir);
      const int col = ilast + pt;
      sA0[e][gf][pt] = gf == 0
          ? radiation_A<T>([&](int d0, int d1) { return BX_U(vr + dir * d1, col + d0); }, pt, sgeo[e][pt], a.cOB_u,
                           ix0, ix1, s_fd)
          : radiation_A<T>([&](int d0, int d1) { return BX_V(vr + dir * d1, col + d0); }, pt, sgeo[e][pt], a.cOB_v,
                           ix0, ix1, s_fd);
    }
  }
  __syncthreads();
  // ---- stage-(s+1) (u, v) on the outer columns of R1; the stage-s running total at the outputs
#pragma unroll
  for (int it = 0; it < NOR; it++) {
    const int t = tid + NT * it;
    if (t < NO0) {
      const int e = t / 10, gf = (t % 10) / 5, pt = t % 5 + 1;
      const T rhs = fma(sgeo[e][pt][3], sk0[e][gf] - sA0[e][gf][0], sA0[e][gf][pt]);
      const T y = syo[e][gf][pt - 1];
      (gf == 0 ? ru1 : rv1)[e][6 + pt] = next_input<T, S0>(rhs, dt, y);
      const int tr = e - NG;
      if (tr >= 0 && tr < TYB)
        srto[tr][gf][pt - 1] =
            PAIR == 1 ? (T)(1.0 / 6.0) * (dt * rhs) : fma((T)(1.0 / 3.0), dt * rhs, srti[tr][gf][pt - 1]);
    }
  }
  __syncthreads();

  // ---- stage s+1 at (ilast, tile rows): the RHS from the stage-(s+1) u on R1 and on the i2 arm
  if (tid < TYB && r0 + tid < g.N1 + NG) {
    const int e = tid + NG;
    T xm[6], xp[6], ym[6], yp[6], zm[6], zp[6];
#pragma unroll
    for (int k = 1; k <= 5; k++) {
      const T *m = sarm[tid][k - 1], *pp = sarm[tid][4 + k];
      zm[k] = next_input<T, S0>(fma(-eta, m[0], m[1]), dt, PAIR == 1 ? m[0] : m[2]);
      zp[k] = next_input<T, S0>(fma(-eta, pp[0], pp[1]), dt, PAIR == 1 ? pp[0] : pp[2]);
      xm[k] = ru1[e][6 - k]; xp[k] = ru1[e][6 + k];
      ym[k] = ru1[e - k][6]; yp[k] = ru1[e + k][6];
    }
    const T K[5] = {skl[tid][0], skl[tid][1], skl[tid][2], skl[tid][3], skl[tid][4]};
    T ku, kv;
    point_rhs<T>(ru1[e][6], rv1[e][6], skl[tid][6], skl[tid][7], K, xm, xp, ym, yp, zm, zp, eta, ku, kv);
    sk1[tid][0] = ku;
    sk1[tid][1] = kv;
  }
  // ---- stage s+1's radiation terms on the tile rows
  if (tid < TYB * 12) {
    const int tr = tid / 12, gf = (tid % 12) / 6, pt = tid % 6, e = tr + NG;
    if (r0 + tr < g.N1 + NG) {
      const T(*R)[12] = gf == 0 ? ru1 : rv1;
      sA1[tr][gf][pt] = radiation_A<T>([&](int d0, int d1) { return R[e + d1][6 + pt + d0]; }, pt, sgeo[e][pt],
                                       gf == 0 ? a.cOB_u : a.cOB_v, ix0, ix1, s_fd);
    }
  }
  __syncthreads();
  // ---- outputs on the outer layers of the tile rows, and their parity images
  if (tid < TYB * 10) {
    const int tr = tid / 10, gf = (tid % 10) / 5, pt = tid % 5 + 1, row = r0 + tr, e = tr + NG;
    if (row < g.N1 + NG) {
      const unsigned p = (unsigned)i2 * s2 + (unsigned)row * s1s + (ilast + pt);
      const T rhs = fma(sgeo[e][pt][3], sk1[tr][gf] - sA1[tr][gf][0], sA1[tr][gf][pt]);
      const T rt = srto[tr][gf][pt - 1], y = syo[e][gf][pt - 1];
      T o;
      if (PAIR == 1) {
        st32(a.rt, gf * nt + p, fma((T)(1.0 / 3.0), dt * rhs, rt));
        o = fma((T)0.5, dt * rhs, y);
      } else {
        o = y + fma((T)(1.0 / 6.0), dt * rhs, rt);
      }
      st32(a.out, gf * nt + p, o);
      for_each_image_in_planes(g, ilast + pt, row, i2, [&](long d) { st32(a.out, gf * nt + (unsigned)d, o); });
    }
  }
}

// One configuration. Without a placement table: the nbc_rows*N2 BC blocks first, then the
// interior blocks in XCD-aware order. With one (ord.n == gridDim.x): workgroup L is logical block
// ord.q[L], the BC blocks being logical 0 .. nb-1.
template <typename T, int PAIR, int TX, int PL>
__global__ void __launch_bounds__(NRPYELL_BLOCK, 3) substep2_kernel(SubstepArgs<T> a, const BlockOrder ord) {
  constexpr int NL = interior2_lds<TX, PL>() > bc2_lds() ? interior2_lds<TX, PL>() : bc2_lds();
  __shared__ T lds[NL];
  __shared__ double s_red[NRPYELL_BLOCK / 64];
  const unsigned nb = (unsigned)(a.nbc_rows * a.g.N2), ni = gridDim.x - nb;
  const bool ordered = ord.n == (int)gridDim.x;
  const unsigned v = ord.at(blockIdx.x);  // unconditional: a scalar load issued with the other arguments
  const unsigned L = ordered ? v : blockIdx.x;
  if (L < nb) {
    const unsigned pln = div_by(L, a.nbc_mul);
    bc2_body<T, PAIR>(a, (int)(L - pln * (unsigned)a.nbc_rows), (int)pln, lds);
    return;
  }
  const unsigned q = ordered ? L - nb : xcd_logical(L - nb, ni), ng = (unsigned)(a.g.N2 / PL);
  const unsigned tile = div_by(q, a.ng_mul);
  interior2_body<T, PAIR, TX, PL>(a, (int)tile, (int)(q - tile * ng), q, ni, lds, s_red);
}

// Many configurations (uneven grids) in one launch, as substep_batched_kernel does for one
// substep: logical block q works for configuration block_cfg[q], as that configuration's block
// q - block_first[cfg] (its BC blocks first, then its interior blocks); a configuration that has
// converged returns at once (its own done flag). Block 0 of a configuration's interior blocks
// of the stages-3+4 launch reduces that configuration's residual partials.
template <int PAIR, int TX, int PL>
__global__ void __launch_bounds__(NRPYELL_BLOCK, 3)
    substep2_batched_kernel(const SubstepArgs<double> *__restrict__ cfgs, const int *__restrict__ block_cfg,
                            const int *__restrict__ block_first) {
  constexpr int NL = interior2_lds<TX, PL>() > bc2_lds() ? interior2_lds<TX, PL>() : bc2_lds();
  __shared__ double lds[NL];
  __shared__ double s_red[NRPYELL_BLOCK / 64];
  const unsigned q = xcd_logical(blockIdx.x, gridDim.x);
  const int c = block_cfg[q];
  const SubstepArgs<double> a = cfgs[c];
  if (*a.done) return;  // leave before issuing any load
  const unsigned local = q - block_first[c], nb = (unsigned)(a.nbc_rows * a.g.N2);
  if (local < nb) {
    const unsigned pln = div_by(local, a.nbc_mul);
    bc2_body<double, PAIR>(a, (int)(local - pln * (unsigned)a.nbc_rows), (int)pln, lds);
    return;
  }
  const unsigned li = local - nb, ng = (unsigned)(a.g.N2 / PL);
  const unsigned tile = div_by(li, a.ng_mul);
  interior2_body<double, PAIR, TX, PL>(a, (int)tile, (int)(li - tile * ng), li, a.ntile0 * a.ntile1 * ng, lds, s_red);
}

template <int PAIR>
static void launch_pair_batched(const SubstepArgs<double> *cfgs, const int *block_cfg, const int *block_first,
                                int nblocks, int tx, int pl, hipStream_t st) {
  dim3 grd(nblocks), blk(NRPYELL_BLOCK);
  if (pl == 2) {
    switch (tx) {
      case 64: hipLaunchKernelGGL((substep2_batched_kernel<PAIR, 64, 2>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
      default: hipLaunchKernelGGL((substep2_batched_kernel<PAIR, 32, 2>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    }
    return;
  }
  switch (tx) {
    case 16: hipLaunchKernelGGL((substep2_batched_kernel<PAIR, 16, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    case 64: hipLaunchKernelGGL((substep2_batched_kernel<PAIR, 64, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
    default: hipLaunchKernelGGL((substep2_batched_kernel<PAIR, 32, 1>), grd, blk, 0, st, cfgs, block_cfg, block_first); break;
  }
}

void launch_substep_pair_batched(const SubstepArgs<double> *cfgs, const int *block_cfg, const int *block_first,
                                 int nblocks, int tx, int pl, int first_stage, hipStream_t st) {
  if (first_stage == 1) launch_pair_batched<1>(cfgs, block_cfg, block_first, nblocks, tx, pl, st);
  else launch_pair_batched<3>(cfgs, block_cfg, block_first, nblocks, tx, pl, st);
}

template <typename T, int PAIR>
static void launch_pair(const SubstepArgs<T> &a, hipStream_t st, const BlockOrder &o) {
  dim3 grd(a.nbc_rows * a.g.N2 + a.ntile0 * a.ntile1 * (a.g.N2 / a.pl)), blk(NRPYELL_BLOCK);
  if (a.pl == 2) {
    switch (a.tx) {
      case 64: hipLaunchKernelGGL((substep2_kernel<T, PAIR, 64, 2>), grd, blk, 0, st, a, o); break;
      default: hipLaunchKernelGGL((substep2_kernel<T, PAIR, 32, 2>), grd, blk, 0, st, a, o); break;
    }
    return;
  }
  switch (a.tx) {
    case 16: hipLaunchKernelGGL((substep2_kernel<T, PAIR, 16, 1>), grd, blk, 0, st, a, o); break;
    case 64: hipLaunchKernelGGL((substep2_kernel<T, PAIR, 64, 1>), grd, blk, 0, st, a, o); break;
    default: hipLaunchKernelGGL((substep2_kernel<T, PAIR, 32, 1>), grd, blk, 0, st, a, o); break;
  }
}

template <typename T>
void launch_substep_pair(const SubstepArgs<T> &a, int first_stage, hipStream_t st, const BlockOrder *ord) {
  static const BlockOrder none{};
  const BlockOrder &o = ord ? *ord : none;
  if (first_stage == 1) launch_pair<T, 1>(a, st, o);
  else launch_pair<T, 3>(a, st, o);
}
template void launch_substep_pair<double>(const SubstepArgs<double> &, int, hipStream_t, const BlockOrder *);
template void launch_substep_pair<float>(const SubstepArgs<float> &, int, hipStream_t, const BlockOrder *);
