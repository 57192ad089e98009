// nrpyell_host.cpp -- host-side setup of libnrpyell_hip.so: grid parameters, the bcstruct
// ghost-zone tables (byte-identical to the reference's), the CFL timestep, and the small
// (i0,i1) tables the kernels read (outer-face geometry, L2 weights, proper spacings).
// Everything here runs once per grid in FP64 with the host libm.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nrpyell_internal.h"

namespace {
constexpr int NG = NRPYELL_NG;

// SinhSymTP radial scale factor pieces as the reference evaluates them:
// f0 = AMAX (exp(x/w) - exp(-x/w)) / (exp(1/w) - exp(-1/w))   (rfm_precompute_rfmstruct_define.c:10)
struct Sinh {
  double A, w, b, D;
  explicit Sinh(const paramstruct *p)
      : A(p->AMAX), w(p->SINHWAA), b(p->bScale), D(exp(1.0 / p->SINHWAA) - exp(-1 / p->SINHWAA)) {}
  double f0(double x) const { return A * (exp(x / w) - exp(-x / w)) / D; }
};

void cart_of_sinhsymtp(const Sinh &s, double x0, double x1, double x2, double X[3]) {
  const double f0 = s.f0(x0);
  X[0] = f0 * sin(x1) * cos(x2);
  X[1] = f0 * sin(x1) * sin(x2);
  X[2] = sqrt(f0 * f0 + s.b * s.b) * cos(x1);
}

bool interior(const paramstruct *p, int i0, int i1, int i2) {
  return i0 >= NG && i0 < p->Nxx_plus_2NGHOSTS0 - NG && i1 >= NG && i1 < p->Nxx_plus_2NGHOSTS1 - NG && i2 >= NG &&
         i2 < p->Nxx_plus_2NGHOSTS2 - NG;
}

// Eigen-coordinate round trip of one point (bcstruct_set_up.c:27-194): SymTP Cartesian map of
// the coordinates handed in, inverse SymTP map, nearest cell-centred index. Returns false where
// the reference aborts ("Cartesian disagreement", relative tolerance 1e-8 in the SinhSymTP map).
bool round_trip(const paramstruct *p, REAL *const xx[3], int i0, int i1, int i2, int out[3]) {
  const double b2 = p->bScale * p->bScale;
  const double x0 = xx[0][i0], x1 = xx[1][i1], x2 = xx[2][i2];
  const double X = x0 * sin(x1) * cos(x2), Y = x0 * sin(x1) * sin(x2), Z = sqrt(b2 + x0 * x0) * cos(x1);
  const double R2 = X * X + Y * Y + Z * Z;
  const double Q = sqrt(-4 * Z * Z * b2 + b2 * b2 + 2 * b2 * R2 + R2 * R2);
  const double sgn = (double)((Z > 0) - (Z < 0));
  const double y[3] = {sqrt(0.5) * sqrt(-b2 + R2 + Q), acos(sqrt(0.5) * sqrt(R2 / b2 - Q / b2 + 1) * sgn), atan2(Y, X)};
  const double dxx[3] = {p->dxx0, p->dxx1, p->dxx2};
  for (int d = 0; d < 3; d++) {
    const double xmin = xx[d][0] - ((double)(0 - NG) + 0.5) * dxx[d];
    out[d] = (int)((y[d] - xmin - 0.5 * dxx[d] + (double)NG * dxx[d]) / dxx[d] + 0.5);
  }
  const Sinh s(p);
  double A[3], B[3];
  cart_of_sinhsymtp(s, x0, x1, x2, A);
  cart_of_sinhsymtp(s, xx[0][out[0]], xx[1][out[1]], xx[2][out[2]], B);
  const double nrm = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) + 1e-15;
  for (int d = 0; d < 3; d++)
    if (fabs(A[d] - B[d]) > 1e-8 * nrm) return false;
  return true;
}

// Signs of the 10 NRPy+ parity types (bcstruct_set_up.c:204-288): dot products of the
// SinhSymTP unit basis vectors at the ghost point and at its image, evaluated on the
// coordinates handed to bcstruct_set_up. NaN (exp overflow on the eigen grid) gives +1.
bool parity_signs(const paramstruct *p, const double a[3], const double c[3], int8_t par[10]) {
  const double iw = 1.0 / p->SINHWAA, b2 = p->bScale * p->bScale;
  const double Dd = exp(iw) - exp(-iw), K = p->AMAX * p->AMAX / (Dd * Dd);
  const double sa = exp(iw * a[0]) - exp(-iw * a[0]), sc = exp(iw * c[0]) - exp(-iw * c[0]);
  const double sin_a = sin(a[1]), sin_c = sin(c[1]);
  const double f0a2 = sa * sa * K, f0c2 = sc * sc * K;
  const double inv = 1 / (sqrt(f0c2 + b2 * sin_c * sin_c) * sqrt(b2 * sin_a * sin_a + f0a2));
  const double radial = inv * sa * sc * K * cos(a[1]) * cos(c[1]);
  const double polar = inv * sin_a * sin_c * sqrt(f0c2 + b2) * sqrt(b2 + f0a2);
  const double ss = sin(a[2]) * sin(c[2]), cc = cos(a[2]) * cos(c[2]);
  const double e0 = radial + ss * polar + polar * cc;
  const double e1 = radial * ss + radial * cc + polar;
  const double e2 = ss + cc;
  const double v[10] = {1, e0, e1, e2, e0 * e0, e0 * e1, e0 * e2, e1 * e1, e1 * e2, e2 * e2};
  bool ok = true;
  for (int k = 0; k < 10; k++) {
    if (fabs(v[k]) < 1 - 1e-8 || fabs(v[k]) > 1 + 1e-8) ok = false;  // NaN passes, as in the reference
    par[k] = v[k] < 0 ? -1 : 1;
  }
  return ok;
}
}  // namespace

GridDims dims_from_params(const paramstruct *p) {
  GridDims g;
  g.N0 = p->Nxx0; g.N1 = p->Nxx1; g.N2 = p->Nxx2;
  g.G0 = p->Nxx_plus_2NGHOSTS0; g.G1 = p->Nxx_plus_2NGHOSTS1; g.G2 = p->Nxx_plus_2NGHOSTS2;
  g.ntot = (long)g.G0 * g.G1 * g.G2;
  g.invdx0 = p->invdx0; g.invdx1 = p->invdx1; g.invdx2 = p->invdx2;
  return g;
}

extern "C" void set_Cparameters_to_default(paramstruct *params) {
  // set_Cparameters_to_default.c:5-86 (NRPy+ parameter defaults)
  memset(params, 0, sizeof(*params));
  params->Nxx0 = 64; params->Nxx1 = 32; params->Nxx2 = 64;
  params->Nxx_plus_2NGHOSTS0 = 70; params->Nxx_plus_2NGHOSTS1 = 38; params->Nxx_plus_2NGHOSTS2 = 70;
  params->xx0 = params->xx1 = params->xx2 = 1e300;
  params->dxx0 = params->dxx1 = params->dxx2 = 0.1;
  params->xxmin0 = params->xxmin1 = params->xxmin2 = 0.1;
  params->xxmax0 = params->xxmax1 = params->xxmax2 = 0.1;
  params->invdx0 = params->invdx1 = params->invdx2 = 1.0;
  double *rfm_scalars[] = {&params->f0_of_xx0, &params->f1_of_xx1, &params->f2_of_xx1, &params->f2_of_xx0_xx1,
                           &params->f3_of_xx0, &params->f4_of_xx2, &params->f0_of_xx0__D0, &params->f0_of_xx0__DD00,
                           &params->f0_of_xx0__DDD000, &params->f1_of_xx1__D1, &params->f1_of_xx1__DD11,
                           &params->f1_of_xx1__DDD111, &params->f2_of_xx1__D1, &params->f2_of_xx1__DD11,
                           &params->f2_of_xx1__DDD111, &params->f2_of_xx0_xx1__D0, &params->f2_of_xx0_xx1__D1,
                           &params->f2_of_xx0_xx1__DD00, &params->f2_of_xx0_xx1__DD11, &params->f3_of_xx0__D0,
                           &params->f3_of_xx0__DD00, &params->f4_of_xx2__D2, &params->f4_of_xx2__DD22};
  for (double *v : rfm_scalars) *v = 1e300;
  params->bScale = 0.5; params->SINHWAA = 0.2; params->AMAX = 10.0;
  params->eta_damping = 1.0; params->time = 0.0;
  params->uu_wavespeed_at_OB = 1.0; params->vv_wavespeed_at_OB = 1.0;
  params->bare_mass_0 = 1.0; params->bare_mass_1 = 0.0;
  params->S0_z = 0.2;
  params->has_outer_boundary = 0;
  params->outer_bc_type = EXTRAPOLATION_OUTER_BCS;
}

extern "C" void set_Nxx_dxx_invdx_params__and__xx(const int EigenCoord, const int Nxx[3], paramstruct *params,
                                                  REAL *xx[3]) {
  // set_Nxx_dxx_invdx_params__and__xx.c:5-53: cell-centred grid, xx[j] = xxmin + (j-NG+1/2) dxx
  // (the reference build evaluates this as one fused multiply-add).
  params->Nxx0 = Nxx[0]; params->Nxx1 = Nxx[1]; params->Nxx2 = Nxx[2];
  params->Nxx_plus_2NGHOSTS0 = Nxx[0] + 2 * NG;
  params->Nxx_plus_2NGHOSTS1 = Nxx[1] + 2 * NG;
  params->Nxx_plus_2NGHOSTS2 = Nxx[2] + 2 * NG;
  const double pi = M_PI;
  const double lo[3] = {0.0, 0.0, -pi};
  const double hi[3] = {EigenCoord == 0 ? 1.0 : params->AMAX, pi, pi};
  double *dxx[3] = {&params->dxx0, &params->dxx1, &params->dxx2};
  double *inv[3] = {&params->invdx0, &params->invdx1, &params->invdx2};
  for (int d = 0; d < 3; d++) {
    *dxx[d] = (hi[d] - lo[d]) / ((REAL)Nxx[d]);
    *inv[d] = 1.0 / *dxx[d];
    const int n = Nxx[d] + 2 * NG;
    xx[d] = (REAL *)malloc(sizeof(REAL) * n);
    for (int j = 0; j < n; j++) xx[d][j] = fma((REAL)(j - NG) + 0.5, *dxx[d], lo[d]);
  }
}

namespace {

// The tables of one grid, as host_bcstruct hands them out.
struct BcTables {
  std::vector<innerpt_bc_struct> inner;
  std::vector<outerpt_bc_struct> outer[NG * 3];
  bc_info_struct info;
};

// bcstruct_set_up.c:355-520 on the host. The per-point work (the SymTP round trip of every
// ghost point and, for inner points, the 10 parity signs) runs on up to 16 threads over i2
// planes into per-point records; the lists are then assembled serially in the reference's
// storage and face order, so the tables are the same bytes a serial pass produces.
int build_bc_tables(const paramstruct *p, REAL *xx[3], BcTables &T, std::string *err) {
  const int G0 = p->Nxx_plus_2NGHOSTS0, G1 = p->Nxx_plus_2NGHOSTS1, G2 = p->Nxx_plus_2NGHOSTS2;
  const long plane = (long)G0 * G1;
  // per point: source index (-1: interior), flags bit0 round-trip failure, bit1 parity failure
  std::vector<int> src((size_t)plane * G2, -1);
  std::vector<unsigned char> flag((size_t)plane * G2, 0);
  std::vector<int8_t> par((size_t)plane * G2 * 10, 0);
  auto work = [&](int i2) {
    for (int i1 = 0; i1 < G1; i1++)
      for (int i0 = 0; i0 < G0; i0++) {
        if (interior(p, i0, i1, i2)) continue;
        const long k = i0 + (long)G0 * i1 + plane * i2;
        int t[3];
        if (!round_trip(p, xx, i0, i1, i2, t)) {
          flag[k] |= 1;
          continue;
        }
        src[k] = t[0] + G0 * (t[1] + G1 * t[2]);
        if (src[k] == (int)k) continue;
        const double a[3] = {xx[0][i0], xx[1][i1], xx[2][i2]}, c[3] = {xx[0][t[0]], xx[1][t[1]], xx[2][t[2]]};
        if (!parity_signs(p, a, c, &par[10 * k])) flag[k] |= 2;
      }
  };
  const int nth = std::max(1, std::min({16, G2, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; t++)
    pool.emplace_back([&, t] {
      for (int i2 = t; i2 < G2; i2 += nth) work(i2);
    });
  for (auto &th : pool) th.join();

  // Step 1 (bcstruct_set_up.c:355-406): inner points in storage order; the first failure in
  // that order is the one reported.
  T.inner.clear();
  memset(&T.info, 0, sizeof(T.info));
  for (long k = 0; k < plane * G2; k++) {
    if (flag[k] & 1) {
      const int i0 = (int)(k % G0), i1 = (int)((k / G0) % G1), i2 = (int)(k / plane);
      if (err) *err = "Error in SinhSymTP coordinate system: Cartesian disagreement at (" + std::to_string(i0) + "," +
                      std::to_string(i1) + "," + std::to_string(i2) + ")";
      return NRPYELL_EGHOSTMAP;
    }
    if (src[k] < 0 || src[k] == (int)k) continue;
    if (flag[k] & 2) {
      if (err) *err = "Parity evaluated to a value not within 8 significant digits of +1 or -1";
      return NRPYELL_EGHOSTMAP;
    }
    innerpt_bc_struct e;
    memset(&e, 0, sizeof(e));
    e.dstpt = (int)k;
    e.srcpt = src[k];
    memcpy(e.parity, &par[10 * k], 10);
    T.inner.push_back(e);
  }
  T.info.num_inner_boundary_points = (int)T.inner.size();

  // Step 2 (bcstruct_set_up.c:411-520): face loop bounds growing one layer at a time; outer
  // lists stored per access slot dirn + 3*which_gz.
  int lo[3] = {NG, NG, NG}, hi[3] = {G0 - NG, G1 - NG, G2 - NG};
  for (int gz = 0; gz < NG; gz++)
    for (int d = 0; d < 3; d++) {
      int *rl = T.info.bc_loop_bounds[gz][2 * d], *rh = T.info.bc_loop_bounds[gz][2 * d + 1];
      for (int e = 0; e < 3; e++) { rl[2 * e] = lo[e]; rl[2 * e + 1] = hi[e]; }
      rl[2 * d] = lo[d] - 1; rl[2 * d + 1] = lo[d];
      lo[d]--;
      for (int e = 0; e < 3; e++) { rh[2 * e] = lo[e]; rh[2 * e + 1] = hi[e]; }
      rh[2 * d] = hi[d]; rh[2 * d + 1] = hi[d] + 1;
      hi[d]++;
    }
  for (int gz = 0; gz < NG; gz++)
    for (int dirn = 0; dirn < 3; dirn++) {
      std::vector<outerpt_bc_struct> &list = T.outer[dirn + 3 * gz];
      list.clear();
      for (int side = 0; side < 2; side++) {
        const int face = 2 * dirn + side;
        const int *r = T.info.bc_loop_bounds[gz][face];
        for (int i2 = r[4]; i2 < r[5]; i2++)
          for (int i1 = r[2]; i1 < r[3]; i1++)
            for (int i0 = r[0]; i0 < r[1]; i0++) {
              const long k = i0 + (long)G0 * i1 + plane * i2;
              if (flag[k] & 1) {
                if (err) *err = "Error in SinhSymTP coordinate system: Cartesian disagreement";
                return NRPYELL_EGHOSTMAP;
              }
              if (src[k] != (int)k) continue;
              outerpt_bc_struct o;
              o.i0 = (short)i0; o.i1 = (short)i1; o.i2 = (short)i2;
              o.FACEX0 = (int8_t)((face == 0) - (face == 1));
              o.FACEX1 = (int8_t)((face == 2) - (face == 3));
              o.FACEX2 = (int8_t)((face == 4) - (face == 5));
              list.push_back(o);
            }
      }
      T.info.num_pure_outer_boundary_points[gz][dirn] = (int)list.size();
    }
  return NRPYELL_OK;
}

// Tables of the last few grids. bcstruct_set_up reads only the grid sizes, AMAX, SINHWAA,
// bScale, dxx and the xx arrays, so a configuration on a grid seen before (the batched solver,
// learned-guess data generation) gets a copy of the same tables without the per-point work.
struct BcCacheEntry {
  int G[3];
  double AMAX, SINHWAA, bScale, dxx[3];
  std::vector<double> xx[3];
  std::shared_ptr<const BcTables> tables;
};
std::mutex g_bc_mu;
std::list<BcCacheEntry> g_bc_cache_host;
constexpr size_t BC_CACHE_MAX = 16;

bool same_grid(const BcCacheEntry &e, const paramstruct *p, REAL *const xx[3]) {
  const int G[3] = {p->Nxx_plus_2NGHOSTS0, p->Nxx_plus_2NGHOSTS1, p->Nxx_plus_2NGHOSTS2};
  const double dxx[3] = {p->dxx0, p->dxx1, p->dxx2};
  for (int d = 0; d < 3; d++)
    if (e.G[d] != G[d] || e.dxx[d] != dxx[d]) return false;
  if (e.AMAX != p->AMAX || e.SINHWAA != p->SINHWAA || e.bScale != p->bScale) return false;
  for (int d = 0; d < 3; d++)
    if (memcmp(e.xx[d].data(), xx[d], sizeof(double) * G[d]) != 0) return false;
  return true;
}

}  // namespace

int host_bcstruct(const paramstruct *p, REAL *xx[3], bc_struct *bc, std::string *err) {
  memset(bc, 0, sizeof(*bc));
  std::shared_ptr<const BcTables> T;
  {
    std::lock_guard<std::mutex> lk(g_bc_mu);
    for (auto it = g_bc_cache_host.begin(); it != g_bc_cache_host.end(); ++it)
      if (same_grid(*it, p, xx)) {
        T = it->tables;
        g_bc_cache_host.splice(g_bc_cache_host.begin(), g_bc_cache_host, it);  // most recent first
        break;
      }
  }
  if (!T) {
    auto built = std::make_shared<BcTables>();
    const int rc = build_bc_tables(p, xx, *built, err);
    if (rc != NRPYELL_OK) return rc;
    T = built;
    BcCacheEntry e;
    e.G[0] = p->Nxx_plus_2NGHOSTS0; e.G[1] = p->Nxx_plus_2NGHOSTS1; e.G[2] = p->Nxx_plus_2NGHOSTS2;
    e.AMAX = p->AMAX; e.SINHWAA = p->SINHWAA; e.bScale = p->bScale;
    e.dxx[0] = p->dxx0; e.dxx[1] = p->dxx1; e.dxx[2] = p->dxx2;
    for (int d = 0; d < 3; d++) e.xx[d].assign(xx[d], xx[d] + e.G[d]);
    e.tables = T;
    std::lock_guard<std::mutex> lk(g_bc_mu);
    g_bc_cache_host.push_front(std::move(e));
    if (g_bc_cache_host.size() > BC_CACHE_MAX) g_bc_cache_host.pop_back();
  }
  // the caller owns (and frees) malloc'ed copies, as with the reference's tables
  const size_t ni = T->inner.size();
  bc->bc_info = T->info;
  bc->inner_bc_array = (innerpt_bc_struct *)malloc(sizeof(innerpt_bc_struct) * (ni ? ni : 1));
  if (ni) memcpy(bc->inner_bc_array, T->inner.data(), sizeof(innerpt_bc_struct) * ni);
  for (int slot = 0; slot < NG * 3; slot++) {
    const size_t no = T->outer[slot].size();
    bc->pure_outer_bc_array[slot] = (outerpt_bc_struct *)malloc(sizeof(outerpt_bc_struct) * (no + 1));
    if (no) memcpy(bc->pure_outer_bc_array[slot], T->outer[slot].data(), sizeof(outerpt_bc_struct) * no);
  }
  return NRPYELL_OK;
}

int ghostmap_verify(const GridDims &g, const bc_struct *bc, std::string *err) {
  const int n = bc->bc_info.num_inner_boundary_points;
  std::vector<int> seen(g.ntot, 0);
  for (int k = 0; k < n; k++) {
    const innerpt_bc_struct &e = bc->inner_bc_array[k];
    const int i0 = e.dstpt % g.G0, i1 = (e.dstpt / g.G0) % g.G1, i2 = e.dstpt / (g.G0 * g.G1);
    int s0, s1, s2;
    ghost_source(g, i0, i1, i2, s0, s1, s2);
    if (idx3(g, s0, s1, s2) != e.srcpt || e.parity[0] != 1) {
      if (err) *err = "ghost point " + std::to_string(e.dstpt) + " maps to " + std::to_string(e.srcpt) +
                      " in the bcstruct table but to " + std::to_string(idx3(g, s0, s1, s2)) + " in the SinhSymTP reflection map";
      return NRPYELL_EGHOSTMAP;
    }
    seen[e.dstpt] = 1;
  }
  // every image the kernels write is an inner point of the table, and every inner point is written once
  long count = 0;
  bool bad = false;
  auto visit = [&](long d) {
    count++;
    if (d < 0 || d >= g.ntot || seen[d] != 1) bad = true;
    else seen[d] = 2;
  };
  for (int i2 = NG; i2 < g.N2 + NG; i2++)
    for (int i1 = NG; i1 < g.N1 + NG; i1++)
      for (int i0 = NG; i0 < g.G0; i0++) {
        if (i0 >= g.N0 + NG) {
          // pure outer points: only those the table does not list as inner
          if (seen[idx3(g, i0, i1, i2)]) continue;
        }
        for_each_image(g, i0, i1, i2, visit);
      }
  if (bad || count != n) {
    if (err) *err = "closed-form image enumeration does not invert the bcstruct table (" + std::to_string(count) +
                    " images for " + std::to_string(n) + " inner points)";
    return NRPYELL_EGHOSTMAP;
  }
  // the interior-plane enumeration of the solver kernels is exactly the interior-plane part
  // of the full one
  std::vector<long> full, part;
  for (int i2 = NG; i2 < g.N2 + NG; i2++)
    for (int i1 = NG; i1 < g.N1 + NG; i1++)
      for (int i0 = NG; i0 < g.G0; i0++) {
        if (i0 >= g.N0 + NG && seen[idx3(g, i0, i1, i2)]) continue;
        full.clear();
        part.clear();
        for_each_image(g, i0, i1, i2, [&](long d) {
          const long j2 = d / ((long)g.G0 * g.G1);
          if (j2 >= NG && j2 < g.N2 + NG) full.push_back(d);
        });
        for_each_image_in_planes(g, i0, i1, i2, [&](long d) { part.push_back(d); });
        std::sort(full.begin(), full.end());
        std::sort(part.begin(), part.end());
        if (full != part) {
          if (err) *err = "interior-plane image enumeration differs from the full one at (" + std::to_string(i0) + "," +
                          std::to_string(i1) + "," + std::to_string(i2) + ")";
          return NRPYELL_EGHOSTMAP;
        }
      }
  return NRPYELL_OK;
}

// Proper spacings (find_timestep.c:19-32, wavespeed_gf_all_points.c:19-36):
//   ds0 = dxx0 f0' f2/f3, ds1 = dxx1 f2, ds2 = dxx2 f0 sin x1.
static void proper_ds(const paramstruct *p, double D, double x0, double x1, double ds[3]) {
  const double iw = 1.0 / p->SINHWAA, ep = exp(iw * x0), em = exp(-iw * x0);
  const double s = sin(x1), f0sq = p->AMAX * p->AMAX * (ep - em) * (ep - em) / (D * D);
  const double f2 = sqrt(p->bScale * p->bScale * s * s + f0sq), AD = p->AMAX / D;
  ds[0] = p->dxx0 * AD * f2 * (iw * ep + iw * em) / sqrt(p->bScale * p->bScale + f0sq);
  ds[1] = p->dxx1 * f2;
  ds[2] = p->dxx2 * s * AD * (ep - em);
}

extern "C" REAL find_timestep(const paramstruct *params, REAL *xx[3], const REAL CFL_FACTOR) {
  // find_timestep.c:5-46. The proper spacing depends on (xx0,xx1) only; the minimum over the
  // interior is taken in the reference's loop order. CFL_FACTOR is not applied (:45).
  (void)CFL_FACTOR;
  const double D = exp(1.0 / params->SINHWAA) - exp(-1.0 / params->SINHWAA);
  double dsmin = 1e38;
  for (int i1 = NG; i1 < NG + params->Nxx1; i1++)
    for (int i0 = NG; i0 < NG + params->Nxx0; i0++) {
      double ds[3];
      proper_ds(params, D, xx[0][i0], xx[1][i1], ds);
      const double m = fmin(fabs(ds[0]), fmin(fabs(ds[1]), fabs(ds[2])));
      if (m < dsmin) dsmin = m;
    }
  return dsmin * 1.0 / 1.0;
}

void host_min_ds_2d(const paramstruct *p, REAL *const xx[3], double *ds2d) {
  const double D = exp(1.0 / p->SINHWAA) - exp(-1.0 / p->SINHWAA);
  const int G0 = p->Nxx_plus_2NGHOSTS0;
  for (int i1 = NG; i1 < NG + p->Nxx1; i1++)
    for (int i0 = NG; i0 < NG + p->Nxx0; i0++) {
      double ds[3];
      proper_ds(p, D, xx[0][i0], xx[1][i1], ds);
      ds2d[i0 + (long)G0 * i1] = ds[0] < (ds[1] < ds[2] ? ds[1] : ds[2]) ? ds[0] : (ds[1] < ds[2] ? ds[1] : ds[2]);
    }
}

// r and (dxx0/dr, dxx1/dr, dxx2/dr) at fixed spherical angles
// (r_and_partial_xi_partial_r_derivs, apply_bcs_outerradiation_and_inner.c:290-321), from the
// SinhSymTP map: r^2 = f0^2 + b^2 cos^2 x1, cos(theta) = f3 cos x1 / r; inverting the 2x2
// Jacobian d(r,cos theta)/d(xx0,xx1) gives dxx0/dr = g1/det, dxx1/dr = -g0/det.
void host_radial_geometry(const paramstruct *p, double x0, double x1, double out[4]) {
  const double w = p->SINHWAA, A = p->AMAX, b = p->bScale;
  const double D = exp(1.0 / w) - exp(-1.0 / w);
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
  out[3] = 0.0;
}

// Per outer layer pt = 0..5 (pt = 0: last interior i0) and i1:
//   [1/r, dxx0/dr, dxx1/dr, (r_0/r_pt)^3].
void host_outer_geometry(const paramstruct *p, REAL *const xx[3], double *geo) {
  const int G1 = p->Nxx_plus_2NGHOSTS1, ilast = p->Nxx0 + NG - 1;
  for (int i1 = 0; i1 < G1; i1++) {
    double o0[4];
    host_radial_geometry(p, xx[0][ilast], xx[1][i1], o0);
    for (int pt = 0; pt < 6; pt++) {
      double o[4];
      host_radial_geometry(p, xx[0][ilast + pt], xx[1][i1], o);
      const double q = o0[0] / o[0];
      double *e = geo + ((long)pt * G1 + i1) * 4;
      e[0] = 1.0 / o[0];
      e[1] = o[1];
      e[2] = o[2];
      e[3] = q * q * q;
    }
  }
}

// L2_norm_of_gf.c:14-35: dV = sqrt(det gamma_hat) dxx0 dxx1 dxx2 inside r < radius, 0 outside;
// sqrt(det gamma_hat) = f0 f0' f2^2 |sin x1| / f3 and r^2 = sin^2 x1 f0^2 + (b^2+f0^2) cos^2 x1.
void host_l2_weights(const paramstruct *p, REAL *const xx[3], double radius, double *w2d, double *volume) {
  const int G0 = p->Nxx_plus_2NGHOSTS0, G1 = p->Nxx_plus_2NGHOSTS1;
  const double w = p->SINHWAA, A = p->AMAX, b = p->bScale;
  const double D = exp(1.0 / w) - exp(-1.0 / w);
  for (long k = 0; k < (long)G0 * G1; k++) w2d[k] = 0.0;
  double V = 0.0;
  for (int i1 = NG; i1 < NG + p->Nxx1; i1++)
    for (int i0 = NG; i0 < NG + p->Nxx0; i0++) {
      const double ep = exp(xx[0][i0] / w), em = exp(-xx[0][i0] / w);
      const double f0 = A * (ep - em) / D, f0p = A * (ep + em) / (w * D);
      const double s = sin(xx[1][i1]), c = cos(xx[1][i1]);
      const double f3sq = f0 * f0 + b * b, f2sq = f0 * f0 + b * b * s * s;
      const double r = sqrt(s * s * f0 * f0 + f3sq * c * c);
      if (r < radius) {
        const double dV = f0 * f0p * f2sq * fabs(s) / sqrt(f3sq) * p->dxx0 * p->dxx1 * p->dxx2;
        w2d[i0 + (long)G0 * i1] = dV;
        V += dV;
      }
    }
  *volume = V * p->Nxx2;  // every i2 plane carries the same weights
}
