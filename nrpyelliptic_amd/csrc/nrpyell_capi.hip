// nrpyell_capi.hip -- the C ABI of libnrpyell_hip.so: the reference driver's function surface
// (NRPy_function_prototypes.h) over HIP, and the resident solver (nrpyell_solver_*).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "nrpyell_internal.h"


namespace {
thread_local std::string g_last_error;
std::mutex g_mu;
}  // namespace

void set_last_error(const std::string &s) { g_last_error = s; }

[[noreturn]] void fatal(const char *what, const std::string &msg) {
  fprintf(stderr, "libnrpyell_hip: %s failed: %s\n", what, msg.c_str());
  exit(1);
}

extern "C" const char *nrpyell_last_error(void) { return g_last_error.c_str(); }

static hipStream_t default_stream() { return 0; }
// drops the cached MoL_step_forward_in_time solvers that use this buffer (defined below)
static void step_cache_forget_buffer(const void *p);

// ---------------------------------------------------------------------------------------
// plumbing
// ---------------------------------------------------------------------------------------
extern "C" int nrpyell_device_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_malloc_device(void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_free_device(void *ptr) {
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}
extern "C" int nrpyell_memset_device(void *ptr, int value, size_t bytes) {
  hipError_t e = hipMemset(ptr, value, bytes);
  if (e != hipSuccess) { set_last_error(hipGetErrorString(e)); return NRPYELL_EHIP; }
  return NRPYELL_OK;
}

template <typename T>
static T *dalloc(size_t n) {
  void *p = nullptr;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  return (T *)p;
}
// Pinned, mapped host memory for a small result that a kernel writes and the host reads after a
// stream synchronisation, without a device-to-host copy; *dev is the address kernels use.
template <typename T>
static T *halloc_mapped(size_t n, T **dev) {
  void *h = nullptr, *d = nullptr;
  HIPCHK(hipHostMalloc(&h, n * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
  memset(h, 0, n * sizeof(T));
  HIPCHK(hipHostGetDevicePointer(&d, h, 0));
  *dev = (T *)d;
  return (T *)h;
}
template <typename T>
static T *upload(const T *h, size_t n) {
  T *d = dalloc<T>(n);
  HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// ---------------------------------------------------------------------------------------
// Reference surface: memory and setup
// ---------------------------------------------------------------------------------------
extern "C" void MoL_malloc_y_n_gfs(const paramstruct *params, MoL_gridfunctions_struct *gf) {
  ref_call("MoL_malloc_y_n_gfs", [&]() {
    const long nt = dims_from_params(params).ntot;
    gf->y_n_gfs = dalloc<double>(NUM_EVOL_GFS * nt);
    gf->diagnostic_output_gfs = gf->y_nplus1_running_total_gfs;  // MoL_malloc_y_n_gfs.c:14
  });
}
extern "C" void MoL_free_memory_y_n_gfs(const paramstruct *params, MoL_gridfunctions_struct *gf) {
  ref_call("MoL_free_memory_y_n_gfs", [&]() {
    (void)params;
    step_cache_forget_buffer(gf->y_n_gfs);
    HIPCHK(hipFree(gf->y_n_gfs));
    gf->y_n_gfs = nullptr;
  });
}
extern "C" void MoL_malloc_non_y_n_gfs(const paramstruct *params, MoL_gridfunctions_struct *gf) {
  ref_call("MoL_malloc_non_y_n_gfs", [&]() {
    const long nt = dims_from_params(params).ntot;
    gf->y_nplus1_running_total_gfs = dalloc<double>(NUM_EVOL_GFS * nt);
    gf->k_odd_gfs = dalloc<double>(NUM_EVOL_GFS * nt);
    gf->k_even_gfs = dalloc<double>(NUM_EVOL_GFS * nt);
    gf->auxevol_gfs = dalloc<double>(NUM_AUXEVOL_GFS * nt);
    // every array starts defined (the reference's malloc'd buffers are fully written before use)
    HIPCHK(hipMemset(gf->y_nplus1_running_total_gfs, 0, NUM_EVOL_GFS * nt * sizeof(double)));
    HIPCHK(hipMemset(gf->k_odd_gfs, 0, NUM_EVOL_GFS * nt * sizeof(double)));
    HIPCHK(hipMemset(gf->k_even_gfs, 0, NUM_EVOL_GFS * nt * sizeof(double)));
    HIPCHK(hipMemset(gf->auxevol_gfs, 0, NUM_AUXEVOL_GFS * nt * sizeof(double)));
    gf->diagnostic_output_gfs = gf->y_nplus1_running_total_gfs;  // MoL_malloc_non_y_n_gfs.c:17
  });
}
extern "C" void MoL_free_memory_non_y_n_gfs(const paramstruct *params, MoL_gridfunctions_struct *gf) {
  ref_call("MoL_free_memory_non_y_n_gfs", [&]() {
    (void)params;
    step_cache_forget_buffer(gf->y_nplus1_running_total_gfs);
    step_cache_forget_buffer(gf->k_odd_gfs);
    step_cache_forget_buffer(gf->k_even_gfs);
    step_cache_forget_buffer(gf->auxevol_gfs);
    HIPCHK(hipFree(gf->y_nplus1_running_total_gfs));
    HIPCHK(hipFree(gf->k_odd_gfs));
    HIPCHK(hipFree(gf->k_even_gfs));
    HIPCHK(hipFree(gf->auxevol_gfs));
    gf->y_nplus1_running_total_gfs = gf->k_odd_gfs = gf->k_even_gfs = gf->auxevol_gfs = nullptr;
    gf->diagnostic_output_gfs = nullptr;
  });
}

extern "C" void rfm_precompute_rfmstruct_malloc(const paramstruct *params, rfm_struct *r) {
  ref_call("rfm_precompute_rfmstruct_malloc", [&]() {
    const long G0 = params->Nxx_plus_2NGHOSTS0, G1 = params->Nxx_plus_2NGHOSTS1;
    r->f0_of_xx0 = dalloc<double>(G0); r->f0_of_xx0__D0 = dalloc<double>(G0);
    r->f0_of_xx0__DD00 = dalloc<double>(G0); r->f0_of_xx0__DDD000 = dalloc<double>(G0);
    r->f1_of_xx1 = dalloc<double>(G1); r->f1_of_xx1__D1 = dalloc<double>(G1); r->f1_of_xx1__DD11 = dalloc<double>(G1);
    r->f2_of_xx0_xx1 = dalloc<double>(G0 * G1); r->f2_of_xx0_xx1__D0 = dalloc<double>(G0 * G1);
    r->f2_of_xx0_xx1__D1 = dalloc<double>(G0 * G1); r->f2_of_xx0_xx1__DD00 = dalloc<double>(G0 * G1);
    r->f2_of_xx0_xx1__DD11 = dalloc<double>(G0 * G1);
    r->f3_of_xx0 = dalloc<double>(G0); r->f3_of_xx0__D0 = dalloc<double>(G0); r->f3_of_xx0__DD00 = dalloc<double>(G0);
  });
}
extern "C" void rfm_precompute_rfmstruct_define(const paramstruct *params, REAL *xx[3], rfm_struct *r) {
  ref_call("rfm_precompute_rfmstruct_define", [&]() {
    const GridDims g = dims_from_params(params);
    double *x0 = upload(xx[0], g.G0), *x1 = upload(xx[1], g.G1);
    launch_rfm(g, params->AMAX, params->SINHWAA, params->bScale, x0, x1, *r, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(x0));
    HIPCHK(hipFree(x1));
  });
}
extern "C" void rfm_precompute_rfmstruct_freemem(const paramstruct *params, rfm_struct *r) {
  ref_call("rfm_precompute_rfmstruct_freemem", [&]() {
    (void)params;
    step_cache_forget_buffer(r->f0_of_xx0);
    double **all[] = {&r->f0_of_xx0, &r->f0_of_xx0__D0, &r->f0_of_xx0__DD00, &r->f0_of_xx0__DDD000, &r->f1_of_xx1,
                      &r->f1_of_xx1__D1, &r->f1_of_xx1__DD11, &r->f2_of_xx0_xx1, &r->f2_of_xx0_xx1__D0,
                      &r->f2_of_xx0_xx1__D1, &r->f2_of_xx0_xx1__DD00, &r->f2_of_xx0_xx1__DD11, &r->f3_of_xx0,
                      &r->f3_of_xx0__D0, &r->f3_of_xx0__DD00};
    for (double **p : all) {
      HIPCHK(hipFree(*p));
      *p = nullptr;
    }
  });
}

extern "C" void initial_guess_all_points(const paramstruct *params, REAL *xx[3], REAL *in_gfs) {
  ref_call("initial_guess_all_points", [&]() {
    // initial_guess_single_point.c:14-15: u = v = 0 at every point, ghosts included
    (void)xx;
    const long nt = dims_from_params(params).ntot;
    HIPCHK(hipMemset(in_gfs, 0, NUM_EVOL_GFS * nt * sizeof(double)));
    HIPCHK(hipDeviceSynchronize());
  });
}

static Punctures punctures_of(const paramstruct *p) {
  Punctures pu;
  pu.m[0] = p->bare_mass_0; pu.m[1] = p->bare_mass_1;
  const double pos[2][3] = {{p->puncture_0_x, p->puncture_0_y, p->puncture_0_z},
                            {p->puncture_1_x, p->puncture_1_y, p->puncture_1_z}};
  const double P[2][3] = {{p->P0_x, p->P0_y, p->P0_z}, {p->P1_x, p->P1_y, p->P1_z}};
  const double S[2][3] = {{p->S0_x, p->S0_y, p->S0_z}, {p->S1_x, p->S1_y, p->S1_z}};
  memcpy(pu.pos, pos, sizeof(pos));
  memcpy(pu.P, P, sizeof(P));
  memcpy(pu.S, S, sizeof(S));
  return pu;
}

extern "C" void auxevol_gfs_all_points(const paramstruct *params, REAL *xx[3], REAL *auxevol_gfs) {
  ref_call("auxevol_gfs_all_points", [&]() {
    const GridDims g = dims_from_params(params);
    double *x0 = upload(xx[0], g.G0), *x1 = upload(xx[1], g.G1), *x2 = upload(xx[2], g.G2);
    launch_auxevol(g, params->AMAX, params->SINHWAA, params->bScale, punctures_of(params), x0, x1, x2, auxevol_gfs,
                   default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(x0)); HIPCHK(hipFree(x1)); HIPCHK(hipFree(x2));
  });
}

extern "C" void wavespeed_gf_all_points(const paramstruct *params, const REAL CFL_FACTOR, const REAL dt, REAL *xx[3],
                                        REAL *auxevol_gfs) {
  ref_call("wavespeed_gf_all_points", [&]() {
    const GridDims g = dims_from_params(params);
    std::vector<double> ds((size_t)g.G0 * g.G1, 0.0);
    host_min_ds_2d(params, xx, ds.data());
    // c = CFL*dsmin/dt (wavespeed_gf_all_points.c:40-41), rounded on the host per (i0,i1)
    for (double &v : ds) v = CFL_FACTOR * v / dt;
    double *d = upload(ds.data(), ds.size());
    launch_wavespeed(g, d, 1.0, dt, auxevol_gfs, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(d));
  });
}

extern "C" REAL compute_wavespeed_at_OB(const paramstruct *params, REAL *auxevol_gfs) {
  return ref_call("compute_wavespeed_at_OB", [&]() -> REAL {
    // compute_wavespeed_at_OB.c:10
    const GridDims g = dims_from_params(params);
    double c = 0.0;
    const long p = idx3(g, g.G0 - NRPYELL_NG - 1, NRPYELL_NG, g.G2 / 2);
    HIPCHK(hipMemcpy(&c, auxevol_gfs + WAVESPEEDGF * g.ntot + p, sizeof(double), hipMemcpyDeviceToHost));
    return c;
  });
}

extern "C" int nrpyell_bcstruct_set_up_checked(const paramstruct *params, REAL *xx[3], bc_struct *bcstruct) {
  std::string err;
  const int rc = host_bcstruct(params, xx, bcstruct, &err);
  if (rc != NRPYELL_OK) set_last_error(err);
  return rc;
}

extern "C" void bcstruct_set_up(const paramstruct *params, REAL *xx[3], bc_struct *bcstruct) {
  std::string err;
  if (host_bcstruct(params, xx, bcstruct, &err) != NRPYELL_OK) {
    fprintf(stderr, "%s\n", err.c_str());  // bcstruct_set_up.c:178-184
    exit(1);
  }
}

extern "C" int nrpyell_ghostmap_check(const paramstruct *params, const bc_struct *bcstruct) {
  std::string err;
  const int rc = ghostmap_verify(dims_from_params(params), bcstruct, &err);
  if (rc != NRPYELL_OK) set_last_error(err);
  return rc;
}

extern "C" void nrpyell_bcstruct_free(bc_struct *bc) {
  step_cache_forget_buffer(bc->inner_bc_array);
  free(bc->inner_bc_array);
  bc->inner_bc_array = nullptr;
  for (int s = 0; s < NGHOSTS * 3; s++) {
    free(bc->pure_outer_bc_array[s]);
    bc->pure_outer_bc_array[s] = nullptr;
  }
}

// ---------------------------------------------------------------------------------------
// Device copies of the host bcstruct tables, cached per table (validated on every use).
// ---------------------------------------------------------------------------------------
struct DeviceBc {
  const innerpt_bc_struct *key;
  int num_inner;
  long ntot;
  std::vector<innerpt_bc_struct> probe;  // sampled host records used to validate the cache
  int *dst = nullptr, *src = nullptr;
  int8_t *par0 = nullptr;
  // radiation lists per slot (dirn + 3*gz), built for one grid: keyed on the values the
  // geometry depends on (AMAX, SINHWAA, bScale, xx0, xx1), not on the xx pointers
  std::vector<double> geo_key;
  RadPoint *rad[NGHOSTS * 3] = {};
  int nrad[NGHOSTS * 3] = {};
  bool rows = false;  // x0 face only, same rows in every layer: one launch for all layers
  // the same points with the Einstein Toolkit thorn's 7-point stencil shifts
  RadPoint *som[NGHOSTS * 3] = {};
  ~DeviceBc() {
    (void)hipFree(dst); (void)hipFree(src); (void)hipFree(par0);
    for (auto *p : rad) (void)hipFree(p);
    for (auto *p : som) (void)hipFree(p);
  }
};

static std::map<const void *, std::unique_ptr<DeviceBc>> g_bc_cache;

static bool bc_matches(const DeviceBc &d, const bc_struct *bc, long ntot) {
  const int n = bc->bc_info.num_inner_boundary_points;
  if (d.key != bc->inner_bc_array || d.num_inner != n || d.ntot != ntot) return false;
  for (size_t k = 0; k < d.probe.size(); k++) {
    const long j = (long)k * (n - 1) / (long)(d.probe.size() > 1 ? d.probe.size() - 1 : 1);
    if (memcmp(&d.probe[k], &bc->inner_bc_array[j], sizeof(innerpt_bc_struct)) != 0) return false;
  }
  return true;
}

static DeviceBc &device_bc(const paramstruct *params, const bc_struct *bc) {
  const GridDims g = dims_from_params(params);
  auto it = g_bc_cache.find(bc->inner_bc_array);
  if (it != g_bc_cache.end() && bc_matches(*it->second, bc, g.ntot)) return *it->second;
  auto d = std::make_unique<DeviceBc>();
  const int n = bc->bc_info.num_inner_boundary_points;
  d->key = bc->inner_bc_array;
  d->num_inner = n;
  d->ntot = g.ntot;
  for (int k = 0; k < 64 && n > 0; k++) d->probe.push_back(bc->inner_bc_array[(long)k * (n - 1) / 63]);
  std::vector<int> hd(n), hs(n);
  std::vector<int8_t> hp(n);
  for (int k = 0; k < n; k++) {
    hd[k] = bc->inner_bc_array[k].dstpt;
    hs[k] = bc->inner_bc_array[k].srcpt;
    hp[k] = bc->inner_bc_array[k].parity[0];  // evol_gf_parity = {0,0} (NRPy_basic_defines.h:199)
  }
  if (n > 0) {
    d->dst = upload(hd.data(), n);
    d->src = upload(hs.data(), n);
    d->par0 = upload(hp.data(), n);
  }
  DeviceBc &ref = *d;
  g_bc_cache[bc->inner_bc_array] = std::move(d);
  return ref;
}

// Clamped stencil offset of compute_partial_r_f (apply_bcs_outerradiation_and_inner.c:348-360).
static int stencil_offset(int idx, int face, int G) {
  int off = face;
  if (idx < 5) off = 5 - idx;
  else if (idx > G - 5 - 1) off = (G - 5 - 1) - idx;
  return off;
}

// Stencil direction of the thorn's 7-point derivative (conformally_flat_BBH_apply_bcs_sommerfeld.h:
// 49-105): one-sided along the face normal and wherever idx + NGHOSTS >= G (backward) or
// idx - NGHOSTS <= 0 (forward); 0 = centred.
static int sommerfeld_shift(int idx, int face, int G) {
  if (abs(face) != 1 && idx + NGHOSTS < G && idx - NGHOSTS > 0) return 0;
  int sh = face;
  if (idx + NGHOSTS >= G) sh = -1;
  if (idx - NGHOSTS <= 0) sh = +1;
  return sh;
}

static void ensure_radiation_lists(DeviceBc &d, const paramstruct *params, const bc_struct *bc, REAL *xx[3]) {
  std::vector<double> key = {params->AMAX, params->SINHWAA, params->bScale};
  key.insert(key.end(), xx[0], xx[0] + params->Nxx_plus_2NGHOSTS0);
  key.insert(key.end(), xx[1], xx[1] + params->Nxx_plus_2NGHOSTS1);
  if (key == d.geo_key) return;
  const GridDims g = dims_from_params(params);
  for (int s = 0; s < NGHOSTS * 3; s++) {
    (void)hipFree(d.rad[s]);
    (void)hipFree(d.som[s]);
    d.rad[s] = d.som[s] = nullptr;
    const int gz = s / 3, dirn = s % 3;
    const int n = bc->bc_info.num_pure_outer_boundary_points[gz][dirn];
    d.nrad[s] = n;
    if (n == 0) continue;
    std::vector<RadPoint> pts(n), spts(n);
    for (int k = 0; k < n; k++) {
      const outerpt_bc_struct &o = bc->pure_outer_bc_array[s][k];
      const int f[3] = {o.FACEX0, o.FACEX1, o.FACEX2};
      const int P[2][3] = {{o.i0, o.i1, o.i2}, {o.i0 + f[0], o.i1 + f[1], o.i2 + f[2]}};
      RadPoint &r = pts[k], &t = spts[k];
      for (int w = 0; w < 2; w++) {
        r.off[w][0] = stencil_offset(P[w][0], f[0], g.G0);
        r.off[w][1] = stencil_offset(P[w][1], f[1], g.G1);
        r.off[w][2] = 0;
        host_radial_geometry(params, xx[0][P[w][0]], xx[1][P[w][1]], r.geo[w]);
        // the thorn evaluates the neighbour's derivative with the outer point's face
        t.off[w][0] = sommerfeld_shift(P[w][0], f[0], g.G0);
        t.off[w][1] = sommerfeld_shift(P[w][1], f[1], g.G1);
        t.off[w][2] = 0;
        for (int c = 0; c < 4; c++) t.geo[w][c] = r.geo[w][c];
      }
      r.p = t.p = idx3(g, P[0][0], P[0][1], P[0][2]);
      r.q = t.q = idx3(g, P[1][0], P[1][1], P[1][2]);
    }
    d.rad[s] = upload(pts.data(), n);
    d.som[s] = upload(spts.data(), n);
  }
  // one-launch form: only x0-face points, and layer gz lists the rows of layer 0 in the same
  // order, each point the outer neighbour of the previous layer's point
  d.rows = d.nrad[0] > 0;
  for (int gz = 0; gz < NGHOSTS && d.rows; gz++) {
    if (d.nrad[3 * gz] != d.nrad[0] || d.nrad[3 * gz + 1] != 0 || d.nrad[3 * gz + 2] != 0) d.rows = false;
    for (int k = 0; k < d.nrad[0] && d.rows && gz > 0; k++) {
      const outerpt_bc_struct &a = bc->pure_outer_bc_array[3 * gz][k], &b = bc->pure_outer_bc_array[3 * (gz - 1)][k];
      if (a.FACEX0 != -1 || b.FACEX0 != -1 || a.i1 != b.i1 || a.i2 != b.i2 || a.i0 != b.i0 + 1) d.rows = false;
    }
  }
  d.geo_key = std::move(key);
}

// ---------------------------------------------------------------------------------------
// Reference surface: operators
// ---------------------------------------------------------------------------------------
extern "C" void rhs_eval(const paramstruct *params, const rfm_struct *rfmstruct, const REAL *auxevol_gfs,
                         const REAL *in_gfs, REAL *rhs_gfs) {
  ref_call("rhs_eval", [&]() {
    const GridDims g = dims_from_params(params);
    launch_rhs_exact(g, rfmstruct, auxevol_gfs, in_gfs, rhs_gfs, params->eta_damping, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  });
}

extern "C" void residual_all_points(const paramstruct *params, const rfm_struct *rfmstruct, const REAL *auxevol_gfs,
                                    const REAL *in_gfs, REAL *residual_gf) {
  ref_call("residual_all_points", [&]() {
    const GridDims g = dims_from_params(params);
    launch_residual_exact(g, rfmstruct, auxevol_gfs, in_gfs, residual_gf, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  });
}

extern "C" void apply_bcs_inner_only(const paramstruct *params, const bc_struct *bcstruct, REAL *gfs) {
  ref_call("apply_bcs_inner_only", [&]() {
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceBc &d = device_bc(params, bcstruct);
    launch_inner_gather(d.ntot, d.num_inner, d.dst, d.src, d.par0, gfs, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  });
}

extern "C" void apply_bcs_outerradiation_and_inner(const paramstruct *params, const bc_struct *bcstruct, REAL *xx[3],
                                                   const REAL custom_wavespeed[NUM_EVOL_GFS],
                                                   const REAL custom_f_infinity[NUM_EVOL_GFS], REAL *gfs,
                                                   REAL *rhs_gfs) {
  ref_call("apply_bcs_outerradiation_and_inner", [&]() {
    std::lock_guard<std::mutex> lk(g_mu);
    const GridDims g = dims_from_params(params);
    DeviceBc &d = device_bc(params, bcstruct);
    ensure_radiation_lists(d, params, bcstruct, xx);
    // innermost layer first; within a layer x0, x1, x2 faces (apply_bcs_outerradiation_and_inner.c:431-458)
    if (d.rows) {
      RadLayers L;
      for (int gz = 0; gz < NGHOSTS; gz++) L.l[gz] = d.rad[3 * gz];
      launch_radiation_rows(g, d.nrad[0], L, gfs, rhs_gfs, custom_wavespeed[0], custom_wavespeed[1],
                            custom_f_infinity[0], custom_f_infinity[1], default_stream());
    } else {
      for (int s = 0; s < NGHOSTS * 3; s++)
        launch_radiation_list(g, d.nrad[s], d.rad[s], gfs, rhs_gfs, custom_wavespeed[0], custom_wavespeed[1],
                              custom_f_infinity[0], custom_f_infinity[1], default_stream());
    }
    launch_inner_gather(d.ntot, d.num_inner, d.dst, d.src, d.par0, rhs_gfs, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  });
}

extern "C" int nrpyell_apply_bcs_sommerfeld(const paramstruct *params, const bc_struct *bcstruct, REAL *xx[3],
                                            const REAL wavespeed[NUM_EVOL_GFS], const REAL f_infinity[NUM_EVOL_GFS],
                                            const REAL radial_falloff_power[NUM_EVOL_GFS], REAL *gfs,
                                            REAL *rhs_gfs) {
  return api_call([&]() -> int {
    if (!params || !bcstruct || !xx || !wavespeed || !f_infinity || !radial_falloff_power || !gfs || !rhs_gfs) {
      set_last_error("nrpyell_apply_bcs_sommerfeld: null argument");
      return NRPYELL_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    const GridDims g = dims_from_params(params);
    DeviceBc &d = device_bc(params, bcstruct);
    ensure_radiation_lists(d, params, bcstruct, xx);
    // innermost layer first: each layer reads the RHS its inner neighbour layer just received
    SommerfeldParams sp;
    for (int k = 0; k < 2; k++) {
      sp.c[k] = wavespeed[k];
      sp.finf[k] = f_infinity[k];
      sp.falloff[k] = radial_falloff_power[k];
    }
    for (int s = 0; s < NGHOSTS * 3; s++) launch_sommerfeld_list(g, d.nrad[s], d.som[s], gfs, rhs_gfs, sp, default_stream());
    launch_inner_gather(d.ntot, d.num_inner, d.dst, d.src, d.par0, rhs_gfs, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NRPYELL_OK;
  });
}

// L2 weight tables cached per (grid, radius, coordinates).
struct L2Weights {
  std::vector<double> key;
  double *w2d = nullptr;
  double *partial = nullptr;
  double *result = nullptr;       // device address of result_host
  double *result_host = nullptr;  // pinned, mapped
  double volume = 0.0;
  ~L2Weights() { (void)hipFree(w2d); (void)hipFree(partial); (void)hipHostFree(result_host); }
};
static std::vector<std::unique_ptr<L2Weights>> g_l2_cache;

static std::vector<double> l2_key(const paramstruct *p, REAL *const xx[3], double radius) {
  std::vector<double> k = {radius, p->AMAX, p->SINHWAA, p->bScale, (double)p->Nxx0, (double)p->Nxx1, (double)p->Nxx2,
                           p->dxx0, p->dxx1, p->dxx2};
  k.insert(k.end(), xx[0], xx[0] + p->Nxx_plus_2NGHOSTS0);
  k.insert(k.end(), xx[1], xx[1] + p->Nxx_plus_2NGHOSTS1);
  return k;
}

static L2Weights &l2_weights(const paramstruct *p, REAL *const xx[3], double radius) {
  std::vector<double> key = l2_key(p, xx, radius);
  for (auto &w : g_l2_cache)
    if (w->key.size() == key.size() && memcmp(w->key.data(), key.data(), key.size() * sizeof(double)) == 0) return *w;
  auto w = std::make_unique<L2Weights>();
  w->key = key;
  const GridDims g = dims_from_params(p);
  std::vector<double> h((size_t)g.G0 * g.G1);
  host_l2_weights(p, xx, radius, h.data(), &w->volume);
  w->w2d = upload(h.data(), h.size());
  const long nblk = (long)((g.N0 + 63) / 64) * ((g.N1 + 3) / 4) * g.N2;
  w->partial = dalloc<double>(nblk);
  w->result_host = halloc_mapped<double>(1, &w->result);
  if (g_l2_cache.size() > 32) g_l2_cache.erase(g_l2_cache.begin());
  g_l2_cache.push_back(std::move(w));
  return *g_l2_cache.back();
}

extern "C" REAL L2_norm_of_gf(const paramstruct *params, const int gf_index, const REAL integration_radius,
                              REAL *xx[3], const REAL *in_gf) {
  return ref_call("L2_norm_of_gf", [&]() -> REAL {
    std::lock_guard<std::mutex> lk(g_mu);
    const GridDims g = dims_from_params(params);
    L2Weights &w = l2_weights(params, xx, integration_radius);
    const long npart = launch_l2_partial(g, in_gf + (long)gf_index * g.ntot, w.w2d, w.partial, default_stream());
    launch_l2_finalize(w.partial, npart, w.volume, nullptr, 0, 0.0, w.result, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(default_stream()));
    return *(volatile double *)w.result_host;
  });
}

// ---------------------------------------------------------------------------------------
// Allocation-free L2 norm. Every device buffer it needs (coordinate values, weight table,
// partial sums, result) lives in `scratch`, one fixed array per GPU that the library holds
// until the process ends. Its device address is looked up once per device; host code uses
// that address, never the symbol itself.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NRPYELL_NOALLOC_SCRATCH_BYTES / sizeof(double)];

#define NRPYELL_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NRPYELL_MAX_DEVICES];

static int scratch_on_current_device(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NRPYELL_MAX_DEVICES) {
    set_last_error("device index beyond the library's table of scratch addresses");
    return NRPYELL_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return NRPYELL_OK;
}

extern "C" int nrpyell_L2_norm_of_gf_noalloc(const paramstruct *params, const int gf_index,
                                             const REAL integration_radius, REAL *xx[3], const REAL *in_gf,
                                             REAL *log10_l2) {
  return api_call([&]() -> int {
    if (!params || !xx || !xx[0] || !xx[1] || !in_gf || !log10_l2 || gf_index < 0) {
      set_last_error("nrpyell_L2_norm_of_gf_noalloc: null argument or negative gf_index");
      return NRPYELL_EINVAL;
    }
    const GridDims g = dims_from_params(params);
    // scratch layout (doubles): xx0 [G0] | xx1 [G1] | weights [G0*G1] | partials | result
    const long nw = (long)g.G0 * g.G1, npart = l2_partial_blocks(g);
    const long need = g.G0 + g.G1 + nw + npart + 1;
    if (need * (long)sizeof(double) > (long)NRPYELL_NOALLOC_SCRATCH_BYTES) {
      set_last_error("nrpyell_L2_norm_of_gf_noalloc: grid needs " + std::to_string(need * sizeof(double)) +
                     " bytes of scratch, more than the " + std::to_string(NRPYELL_NOALLOC_SCRATCH_BYTES) +
                     " it has; use L2_norm_of_gf");
      return NRPYELL_EINVAL;
    }
    double *sp = nullptr;
    const int rc = scratch_on_current_device(&sp);
    if (rc != NRPYELL_OK) return rc;
    double *d_xx0 = sp, *d_xx1 = d_xx0 + g.G0, *d_w = d_xx1 + g.G1, *d_part = d_w + nw, *d_out = d_part + npart;
    const hipStream_t st = default_stream();
    HIPCHK(hipMemcpyAsync(d_xx0, xx[0], g.G0 * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_xx1, xx[1], g.G1 * sizeof(double), hipMemcpyHostToDevice, st));
    const L2WeightParams q = {params->AMAX, params->SINHWAA, params->bScale, params->dxx0,
                              params->dxx1, params->dxx2, integration_radius};
    launch_l2_weights(g, q, d_xx0, d_xx1, d_w, st);
    const long n = launch_l2_partial(g, in_gf + (long)gf_index * g.ntot, d_w, d_part, st);
    launch_l2_finalize_weights(d_part, n, d_w, nw, g.N2, d_out, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(log10_l2, d_out, sizeof(double), hipMemcpyDeviceToHost));
    return NRPYELL_OK;
  });
}

// gridfunction_z_axis.c:6-52: z and gf along the +z axis (i1 = NG), -z axis (i1 = G1-NG-1)
// and the i0 = NG line, at i2 = G2/2, "%.17e %.17e" per line.
extern "C" void gridfunction_z_axis(const paramstruct *params, const int gf_index, REAL *xx[3], const REAL *in_gf,
                                    FILE *outfile) {
  ref_call("gridfunction_z_axis", [&]() {
    const GridDims g = dims_from_params(params);
    std::vector<double> h(g.ntot);
    HIPCHK(hipMemcpy(h.data(), in_gf + (long)gf_index * g.ntot, g.ntot * sizeof(double), hipMemcpyDeviceToHost));
    const double A = params->AMAX, b = params->bScale;
    // z = sqrt(AMAX^2 sinh^2 / sinh^2 + b^2) cos x1, evaluated in the reference's operation order
    // (its GCC build fuses the "+ b^2" into one fma), so the text output is identical
    const double iw = 1.0 / params->SINHWAA;
    const double D = exp(iw) - exp(-iw);
    auto zc = [&](double x0, double x1) {
      const double s = exp(iw * x0) - exp(-iw * x0);
      return sqrt(fma(b, b, (A * A) * (s * s) / (D * D))) * cos(x1);
    };
    const int i2 = g.G2 / 2;
    for (int pass = 0; pass < 2; pass++) {
      const int i1 = pass == 0 ? NRPYELL_NG : g.G1 - NRPYELL_NG - 1;
      for (int i0 = NRPYELL_NG; i0 < g.G0 - NRPYELL_NG; i0++)
        fprintf(outfile, "%.17e %.17e\n", zc(xx[0][i0], xx[1][i1]), h[idx3(g, i0, i1, i2)]);
    }
    for (int i1 = NRPYELL_NG + 1; i1 < g.G1 - NRPYELL_NG - 1; i1++)
      fprintf(outfile, "%.17e %.17e\n", zc(xx[0][NRPYELL_NG], xx[1][i1]), h[idx3(g, NRPYELL_NG, i1, i2)]);
  });
}

// gridfunction_xz_plane.c:6-49: x, z, gf on the xz plane: the x < 0 half at i2 = NGHOSTS,
// then the x > 0 half at i2 = G2/2, "%.17e %.17e %.17e" per line.
extern "C" void gridfunction_xz_plane(const paramstruct *params, const int gf_index, REAL *xx[3], const REAL *in_gf,
                                      FILE *outfile) {
  ref_call("gridfunction_xz_plane", [&]() {
    const GridDims g = dims_from_params(params);
    std::vector<double> h(g.ntot);
    HIPCHK(hipMemcpy(h.data(), in_gf + (long)gf_index * g.ntot, g.ntot * sizeof(double), hipMemcpyDeviceToHost));
    const double A = params->AMAX, b = params->bScale;
    const double iw = 1.0 / params->SINHWAA;
    const double D = exp(iw) - exp(-iw);
    const int i2s[2] = {NRPYELL_NG, g.G2 / 2};
    for (int k = 0; k < 2; k++) {
      const int i2 = i2s[k];
      for (int i1 = NRPYELL_NG; i1 < g.G1 - NRPYELL_NG; i1++)
        for (int i0 = NRPYELL_NG; i0 < g.G0 - NRPYELL_NG; i0++) {
          const double s = exp(iw * xx[0][i0]) - exp(-iw * xx[0][i0]);
          const double x = A * s * sin(xx[1][i1]) * cos(xx[2][i2]) / D;
          const double z = sqrt(fma(b, b, (A * A) * (s * s) / (D * D))) * cos(xx[1][i1]);
          fprintf(outfile, "%.17e %.17e %.17e\n", x, z, h[idx3(g, i0, i1, i2)]);
        }
    }
  });
}

// print_puncture_parameters.c:6-37: the same report on stdout.
extern "C" void print_puncture_parameters(const paramstruct *p) {
  const struct { const char *name; double v; } rows[] = {
      {"bScale      ", p->bScale}, {"SINHWAA     ", p->SINHWAA}, {"AMAX        ", p->AMAX},
      {"eta_damping ", p->eta_damping}, {"bare_mass_0 ", p->bare_mass_0}, {"bare_mass_1 ", p->bare_mass_1},
      {"puncture_0_x", p->puncture_0_x}, {"puncture_0_y", p->puncture_0_y}, {"puncture_0_z", p->puncture_0_z},
      {"puncture_1_x", p->puncture_1_x}, {"puncture_1_y", p->puncture_1_y}, {"puncture_1_z", p->puncture_1_z},
      {"P0_x        ", p->P0_x}, {"P0_y        ", p->P0_y}, {"P0_z        ", p->P0_z},
      {"P1_x        ", p->P1_x}, {"P1_y        ", p->P1_y}, {"P1_z        ", p->P1_z},
      {"S0_x        ", p->S0_x}, {"S0_y        ", p->S0_y}, {"S0_z        ", p->S0_z},
      {"S1_x        ", p->S1_x}, {"S1_y        ", p->S1_y}, {"S1_z        ", p->S1_z}};
  for (const auto &r : rows) printf("%s = %lf\n", r.name, r.v);
  printf("Nxx0         = %d\n", p->Nxx0);
  printf("Nxx1         = %d\n", p->Nxx1);
  printf("Nxx2         = %d\n", p->Nxx2);
  printf("NGHOSTS      = %d\n", NGHOSTS);
}

// ---------------------------------------------------------------------------------------
// Resident solver
// ---------------------------------------------------------------------------------------
struct nrpyell_solver {
  griddata_struct *gd = nullptr;
  GridDims g{};
  double dt = 0.0;
  nrpyell_solve_options opt{};
  hipStream_t stream = nullptr;
  int tx = 32, pl = 1, ntile0 = 0, ntile1 = 0;
  bool pair_path = false;  // FP64 steps as two launches of two substeps each
  int ybuf = 0;            // pair path: y_n currently in y_n_gfs (0) or k_even_gfs (1)
  double *coef64 = nullptr;   // metric coefficients (residual kernel)
  double *kc64 = nullptr;     // per (i0,i1): c^2 x 5 coefficients, L2 weight / c^4 (substep kernel)
  float *kc32 = nullptr;
  double *src64 = nullptr;    // per point: psi, 0.125 c^2 AA^2 (substep kernel)
  double *geo64 = nullptr;
  float *geo32 = nullptr;
  double *w2d = nullptr;
  double volume = 0.0;
  double *partial = nullptr;
  double *l2_out = nullptr;
  SolverDeviceState *dstate = nullptr;
  SolverDeviceState *hstate = nullptr;  // pinned
  // FP32 shadow state
  float *y32 = nullptr, *ko32 = nullptr, *ke32 = nullptr, *rt32 = nullptr, *src32 = nullptr;
  hipGraphExec_t graph64 = nullptr, graph32 = nullptr;
  hipGraphExec_t graph_step = nullptr;  // the drop-in MoL_step_forward_in_time call
  BlockOrder order{};                    // placement of the two-substep kernel's blocks (n == 0: none)
  // what a captured graph bakes into its kernel arguments; a graph is recaptured when any of it
  // changes between solves (eta_damping or the outer-boundary wave speeds edited in the params)
  struct GraphKey {
    double eta = 0, dt = 0, cu = 0, cv = 0, target = 0;
    int steps = 0;
    bool same(const GraphKey &o) const {
      return eta == o.eta && dt == o.dt && cu == o.cu && cv == o.cv && target == o.target && steps == o.steps;
    }
  };
  GraphKey key64, key32, key_step;
  // residual history of the FP32 phase of the last mixed-precision solve (its checks count FP32
  // steps); the FP64 phase's entries follow it, offset by fp32_offset steps
  std::vector<int> hist32_step;
  std::vector<double> hist32_log10;
  int fp32_offset = 0;
  double *cached_aux = nullptr;
  int *bad = nullptr;         // set by the fold kernel when the wave speed varies along xx2
  int *bad_host = nullptr;    // the same flag as the host reads it (pinned, mapped; bad is its device address)
  ~nrpyell_solver() {
    if (graph64) (void)hipGraphExecDestroy(graph64);
    if (graph32) (void)hipGraphExecDestroy(graph32);
    if (graph_step) (void)hipGraphExecDestroy(graph_step);
    (void)hipFree(coef64); (void)hipFree(kc64); (void)hipFree(kc32); (void)hipFree(src64);
    (void)hipFree(geo64); (void)hipFree(geo32); (void)hipFree(w2d); (void)hipFree(partial);
    (void)hipFree(l2_out); (void)hipFree(dstate); (void)hipHostFree(bad_host);
    if (hstate) (void)hipHostFree(hstate);
    (void)hipFree(y32); (void)hipFree(ko32); (void)hipFree(ke32); (void)hipFree(rt32); (void)hipFree(src32);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  SubstepArgs<T> args(int stage, T *yn, T *ko, T *ke, T *rt, const T *src, const T *coef, const T *geo) const {
    SubstepArgs<T> a;
    a.g = g;
    const MoL_gridfunctions_struct &gf = gd->gridfuncs;
    (void)gf;
    switch (stage) {
      case 1: a.in = yn; a.out = ko; break;
      case 2: a.in = ko; a.out = ke; break;
      case 3: a.in = ke; a.out = ko; break;
      default: a.in = ko; a.out = yn; break;
    }
    a.yn = yn;
    a.rt = rt;
    a.src = src;
    a.coef = coef;
    a.geo = geo;
    a.bc_kind = opt.outer_bc;
    const double c_default = opt.outer_bc == NRPYELL_OUTER_BC_SOMMERFELD6 ? NRPYELL_ET_SOMMERFELD_SPEED : 0.0;
    const double c_fixed = opt.outer_bc_speed > 0 ? opt.outer_bc_speed : c_default;
    a.cOB_u = (T)(c_fixed > 0 ? c_fixed : gd->params.uu_wavespeed_at_OB);
    a.cOB_v = (T)(c_fixed > 0 ? c_fixed : gd->params.vv_wavespeed_at_OB);
    a.eta = (T)gd->params.eta_damping;
    a.dt = (T)dt;
    a.done = &dstate->done;
    a.tx = tx;
    a.pl = pl;
    a.nbc_rows = (g.N1 + NRPYELL_BC_ROWS - 1) / NRPYELL_BC_ROWS;
    a.ntile0 = ntile0;
    a.ntile1 = ntile1;
    a.i0_first = NRPYELL_NG + g.N0 - ntile0 * tx;
    a.ng_mul = div_mul((unsigned)(g.N2 / pl));
    a.nt0_mul = div_mul((unsigned)ntile0);
    a.nbc_mul = div_mul((unsigned)a.nbc_rows);
    a.l2_partial = nullptr;
    a.state = nullptr;
    a.l2_volume = 0.0;
    a.l2_target = 0.0;
    return a;
  }

  // The pair path keeps y_n in one of two arrays: y_n_gfs or k_even_gfs (which the pair
  // kernels do not otherwise use); each step writes the other one.
  double *ycur(int j) const { return j == 0 ? gd->gridfuncs.y_n_gfs : gd->gridfuncs.k_even_gfs; }
  // One FP64 RK4 step as two launches of two substeps, from y in ycur(j) to ycur(1 - j); with
  // fuse_check the first launch reduces the residual of y_n into per-block partials and the
  // second one's block 0 updates the convergence state.
  void step64_pair(hipStream_t st, int j, bool fuse_check) {
    MoL_gridfunctions_struct &f = gd->gridfuncs;
    SubstepArgs<double> a = args<double>(1, ycur(j), f.k_odd_gfs, nullptr, f.y_nplus1_running_total_gfs, src64,
                                         kc64, geo64);
    a.in = ycur(j);
    a.yn = ycur(j);
    a.out = f.k_odd_gfs;
    pair_layout(a);
    if (fuse_check) {
      a.l2_partial = partial;
      a.state = dstate;
      a.l2_volume = volume;
      a.l2_target = opt.log10_target;
    }
    launch_substep_pair(a, 1, st, &order);
    a.in = f.k_odd_gfs;
    a.yn = ycur(j);
    a.out = ycur(1 - j);
    launch_substep_pair(a, 3, st, &order);
  }
  // the pair kernel's BC blocks cover NRPYELL_PAIR_BC_ROWS rows each
  template <typename T>
  void pair_layout(SubstepArgs<T> &a) const {
    a.nbc_rows = (g.N1 + NRPYELL_PAIR_BC_ROWS - 1) / NRPYELL_PAIR_BC_ROWS;
    a.nbc_mul = div_mul((unsigned)a.nbc_rows);
  }
  // RK4 steps in a graph: the pair path needs an even count (y_n back in y_n_gfs)
  int graph_steps64() const { return pair_path ? opt.steps_per_graph + (opt.steps_per_graph & 1) : opt.steps_per_graph; }
  // ... and in the FP32 phase also a multiple of check_every (steps_per_graph is one)
  int graph_steps32() const { return pair_path && (opt.steps_per_graph & 1) ? 2 * opt.steps_per_graph : opt.steps_per_graph; }
  // The FP32 state of the pair path alternates between y32 and ke32 like the FP64 one.
  float *ycur32(int j) const { return j == 0 ? y32 : ke32; }
  // One FP32 RK4 step in two launches, from ycur32(j) to ycur32(1 - j) (no fused residual:
  // the FP32 phase checks the residual in FP64 every check_every steps).
  void step32_pair(hipStream_t st, int j) {
    SubstepArgs<float> a = args<float>(1, ycur32(j), ko32, nullptr, rt32, src32, kc32, geo32);
    a.in = ycur32(j);
    a.yn = ycur32(j);
    a.out = ko32;
    pair_layout(a);
    launch_substep_pair(a, 1, st, &order);
    a.in = ko32;
    a.yn = ycur32(j);
    a.out = ycur32(1 - j);
    launch_substep_pair(a, 3, st, &order);
  }
  // One FP64 RK4 step; with fuse_check the stage-1 kernel also reduces the residual L2 of
  // y_n (before the step) into the convergence state.
  void step64(hipStream_t st, bool fuse_check = false) {
    MoL_gridfunctions_struct &f = gd->gridfuncs;
    for (int s = 1; s <= 4; s++) {
      SubstepArgs<double> a = args<double>(s, f.y_n_gfs, f.k_odd_gfs, f.k_even_gfs, f.y_nplus1_running_total_gfs,
                                           src64, kc64, geo64);
      if (s <= 2 && fuse_check) {  // stage 1 writes the partials, stage 2 reduces them
        a.l2_partial = partial;
        a.state = dstate;
        a.l2_volume = volume;
        a.l2_target = opt.log10_target;
      }
      launch_substep<double>(a, s, st);
    }
  }
  GraphKey graph_key(bool fp32, double target) const {
    GraphKey k;
    if (fp32) {
      const SubstepArgs<float> a = args<float>(1, y32, ko32, ke32, rt32, src32, kc32, geo32);
      k.eta = a.eta; k.dt = a.dt; k.cu = a.cOB_u; k.cv = a.cOB_v;
    } else {
      const SubstepArgs<double> a = args<double>(1, nullptr, nullptr, nullptr, nullptr, src64, kc64, geo64);
      k.eta = a.eta; k.dt = a.dt; k.cu = a.cOB_u; k.cv = a.cOB_v;
    }
    k.target = target;
    k.steps = opt.steps_per_graph;
    return k;
  }
  // Metric coefficients from the rfm arrays and the c^2-folded coefficient and source tables
  // from auxevol: the only reads of those arrays. Raises `bad` (cleared first) when the wave
  // speed varies along xx2.
  // also_clear: another flag to zero with the wave-speed check's (the drop-in step's done flag)
  // (the flags are zeroed by a kernel, not a memset: this runs inside the drop-in step's graph)
  void refresh_tables(hipStream_t st, int *also_clear = nullptr) {
    launch_tables(g, &gd->rfmstruct, gd->gridfuncs.auxevol_gfs, w2d, coef64, kc64, kc32, src64, bad, also_clear, st);
    HIPCHK(hipGetLastError());
  }
  // The fused kernels keep only the interior phi planes current; this fills the ghost planes
  // of y_n from them before the state is handed back.
  void fill_ghosts(hipStream_t st) { launch_fill_ghost_planes<double>(g, gd->gridfuncs.y_n_gfs, 2, st); }
  // y_n_gfs handed back after steps that left the state in array j (pair path: 1 = k_even_gfs,
  // copied back with its ghost planes filled in the same pass; 0 = y_n_gfs, ghost planes filled)
  void finish(hipStream_t st, int j) {
    if (j == 1)
      launch_settle_fill(g, gd->gridfuncs.k_even_gfs, gd->gridfuncs.y_n_gfs, 2, st);
    else
      fill_ghosts(st);
  }
  void step32(hipStream_t st) {
    for (int s = 1; s <= 4; s++) launch_substep<float>(args<float>(s, y32, ko32, ke32, rt32, src32, kc32, geo32), s, st);
  }
  // residual of the current y_n (FP64, or the FP32 shadow in u32, y32 by default) and a
  // convergence-state update
  void check(hipStream_t st, bool from32, int steps_per_check, double target, const float *u32 = nullptr) {
    const void *u = from32 ? (const void *)(u32 ? u32 : y32) : (const void *)gd->gridfuncs.y_n_gfs;
    const long npart = launch_residual_l2(g, u, from32, gd->gridfuncs.auxevol_gfs, coef64, w2d, partial,
                                          &dstate->done, nullptr, st);
    launch_l2_finalize(partial, npart, volume, dstate, steps_per_check, target, l2_out, st);
  }
};

// Placement of the two-substep kernel's blocks (BlockOrder). Workgroups are dealt to the XCDs
// round robin (workgroup L to XCD L % 8) and within an XCD to its CUs round robin (its k-th
// workgroup to CU k % P, P = CUs per XCD), as the hardware-id register shows on MI355X. When the
// whole launch is resident at once, each CU's time is set by the blocks it holds, and the blocks
// differ: per-block timestamps at 128x128x16 give BC blocks and last interior tiles (radiation
// terms, outer layers) the longest lives, first tiles the shortest; the kernel's own order happened
// to put the last tile columns on the CUs that also hold BC blocks. Each XCD keeps the blocks the
// kernel's own order gives it (whole tile columns: L2 reuse of the i2 arms and i1 halos). Its BC
// blocks keep the first positions (dispatched first, one per CU); the interior blocks go heaviest
// first to the CU with the least load and a free position, then pairwise swaps of interior blocks
// lower the most loaded CU while they can. Relative weights: BC 1.25, last tile 1.1, other tiles
// 1.0, first tile 0.85. A CU's blocks are then dispatched last tile first, first tile next, other
// tiles after: the order matters as much as the split, measured at 128x128x16 where this gives
// the CUs holding three blocks a BC block, a first tile and another tile, in that order, and the
// CUs holding two a last tile and another. Placements measured there (us per RK4 step): the
// kernel's own order 21.5; BC blocks placed like interior blocks, on the CUs that hold two, 20.6;
// the same split as here with the first tile after the other tile 21.3; a weight of +0.15 for
// tiles on the first or last i1 row 20.3-20.45; this one 20.05-20.1.
// plan_block_order fills `order` for a launch of nb BC blocks and ntile0 x ntile1 x ng interior
// blocks on a GPU of ncu CUs (order.n = 0: the kernel's own order).
static void plan_block_order(int ncu, int nb, int ntile0, int ntile1, int ng, BlockOrder &order) {
  order = BlockOrder{};
  const int ni = ntile0 * ntile1 * ng, n = nb + ni;
  const int P = ncu / 8;
  // one round of resident blocks (3 per CU: __launch_bounds__(256, 3), 37 KB of LDS each)
  if (ncu % 8 != 0 || P <= 0 || n > NRPYELL_ORDER_MAX || n > 3 * ncu || nb % 8 || ni % 8) return;
  const int per = ni / 8, nx = n / 8;
  struct Item {
    double w;
    int logical, rank;  // rank: dispatch order within a CU (0 last tile, 1 first tile, 2 other)
  };
  const int nbx = nb / 8;  // BC blocks per XCD
  if (nbx > P) return;
  for (int x = 0; x < 8; x++) {
    std::vector<Item> mem;
    for (int q = x * per; q < (x + 1) * per; q++) {
      const int t0 = (q / ng) % ntile0;
      const int rank = t0 == ntile0 - 1 ? 0 : t0 == 0 ? 1 : 2;
      mem.push_back({rank == 0 ? 1.1 : rank == 1 ? 0.85 : 1.0, nb + q, rank});
    }
    std::stable_sort(mem.begin(), mem.end(), [](const Item &a, const Item &b) { return a.w > b.w; });
    std::vector<std::vector<Item>> on(P);
    std::vector<double> load(P, 0.0);
    std::vector<int> cap(P, 0);
    for (int j = nbx; j < nx; j++) cap[j % P]++;  // positions left after the BC blocks
    for (int k = 0; k < nbx; k++) {
      load[k] = 1.25;                                // position k, CU k
      order.set(8 * k + x, (unsigned)(8 * k + x));  // BC block 8 k + x: its own order
    }
    for (const Item &m : mem) {
      int best = -1;
      for (int c = 0; c < P; c++)
        if ((int)on[c].size() < cap[c] &&
            (best < 0 || load[c] < load[best] || (load[c] == load[best] && cap[c] < cap[best])))
          best = c;
      on[best].push_back(m);
      load[best] += m.w;
    }
    for (int it = 0; it < 4096; it++) {
      const int cm = (int)(std::max_element(load.begin(), load.end()) - load.begin());
      double gain = 1e-9;
      int bi = -1, bc = -1, bj = -1;
      for (int i = 0; i < (int)on[cm].size(); i++)
        for (int c = 0; c < P; c++)
          for (int j = 0; c != cm && j < (int)on[c].size(); j++) {
            const double d = on[cm][i].w - on[c][j].w;
            const double g = load[cm] - std::max(load[cm] - d, load[c] + d);
            if (g > gain) { gain = g; bi = i; bc = c; bj = j; }
          }
      if (bi < 0) break;
      const double d = on[cm][bi].w - on[bc][bj].w;
      std::swap(on[cm][bi], on[bc][bj]);
      load[cm] -= d;
      load[bc] += d;
    }
    for (int c = 0; c < P; c++) {
      std::stable_sort(on[c].begin(), on[c].end(), [](const Item &a, const Item &b) {
        return a.rank != b.rank ? a.rank < b.rank : a.logical < b.logical;
      });
      int k = 0;
      for (int j = c < nbx ? c + P : c; j < nx; j += P) order.set(8 * j + x, (unsigned)on[c][k++].logical);
    }
  }
  order.n = n;
}

static void pair_block_order(nrpyell_solver *s) {
  s->order = BlockOrder{};
  if (!s->pair_path) return;
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int nbc_rows = (s->g.N1 + NRPYELL_PAIR_BC_ROWS - 1) / NRPYELL_PAIR_BC_ROWS;
  plan_block_order(ncu, nbc_rows * s->g.N2, s->ntile0, s->ntile1, s->g.N2 / s->pl, s->order);
}

extern "C" int nrpyell_block_order_plan(int ncu, int N0, int N1, int N2, int tile_i0, int planes, unsigned short *out,
                                        int max_entries) {
  if (ncu <= 0 || N0 <= 0 || N1 <= 0 || N2 <= 0 || (tile_i0 != 16 && tile_i0 != 32 && tile_i0 != 64) ||
      !(planes == 1 || (planes == 2 && tile_i0 >= 32)) || N2 % planes != 0 || max_entries < 0) {
    set_last_error("nrpyell_block_order_plan: invalid arguments");
    return NRPYELL_EINVAL;
  }
  const int ty = NRPYELL_BLOCK / tile_i0;
  BlockOrder o;
  plan_block_order(ncu, (N1 + NRPYELL_PAIR_BC_ROWS - 1) / NRPYELL_PAIR_BC_ROWS * N2, (N0 + tile_i0 - 1) / tile_i0,
                   (N1 + ty - 1) / ty, N2 / planes, o);
  const int n = !out ? 0 : o.n < max_entries ? o.n : max_entries;
  for (int i = 0; i < n; i++) out[i] = (unsigned short)o.at(i);
  return o.n;
}

extern "C" int nrpyell_solver_block_order(const nrpyell_solver *s, unsigned short *out, int max_entries) {
  if (!s) return 0;
  const int n = !out ? 0 : s->order.n < max_entries ? s->order.n : max_entries;
  for (int i = 0; i < n; i++) out[i] = (unsigned short)s->order.at(i);
  return s->order.n;
}

static void reset_state(nrpyell_solver *s) {
  s->hist32_step.clear();
  s->hist32_log10.clear();
  s->fp32_offset = 0;
  memset(s->hstate, 0, sizeof(SolverDeviceState));
  if (s->opt.stop_rule == NRPYELL_STOP_ET_THORN) {
    s->hstate->rule_every = s->opt.check_every;
    s->hstate->rule_last = s->opt.max_steps + 1;
  }
  HIPCHK(hipMemcpyAsync(s->dstate, s->hstate, sizeof(SolverDeviceState), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
}

static void read_state(nrpyell_solver *s) {
  HIPCHK(hipMemcpyAsync(s->hstate, s->dstate, sizeof(SolverDeviceState), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
}

extern "C" void nrpyell_default_options(nrpyell_solve_options *o) {
  o->precision = 64;
  o->check_every = 10;
  o->steps_per_graph = 50;
  o->log10_target = -10.0;
  o->max_steps = 200000;
  o->fp32_switch_log10 = -6.0;
  o->fp32_plateau_steps = 200;
  o->integration_radius = 100.0;  // main.c:150
  o->tile_i0 = 0;
  o->planes = 0;
  o->outer_bc = NRPYELL_OUTER_BC_RADIATION;
  o->outer_bc_speed = 0.0;
  o->stop_rule = NRPYELL_STOP_EVERY_STEP;
  o->substeps_per_launch = 0;
}

extern "C" int nrpyell_solver_create(griddata_struct *gd, REAL dt, const nrpyell_solve_options *opt,
                                     nrpyell_solver **out) {
  return api_call([&]() -> int {
    std::lock_guard<std::mutex> lk(g_mu);
    auto s = std::make_unique<nrpyell_solver>();
    s->gd = gd;
    s->g = dims_from_params(&gd->params);
    s->dt = dt;
    if (opt) s->opt = *opt; else nrpyell_default_options(&s->opt);
    if (s->opt.stop_rule != NRPYELL_STOP_EVERY_STEP && s->opt.stop_rule != NRPYELL_STOP_ET_THORN) {
      set_last_error("stop_rule must be NRPYELL_STOP_EVERY_STEP or NRPYELL_STOP_ET_THORN");
      return NRPYELL_EINVAL;
    }
    if (s->opt.stop_rule == NRPYELL_STOP_ET_THORN && (s->opt.precision != 64 || s->opt.max_steps < 0)) {
      set_last_error("the thorn's stopping rule needs precision 64 and max_steps >= 0");
      return NRPYELL_EINVAL;
    }
    if (s->opt.check_every < 1) s->opt.check_every = 1;
    if (s->opt.steps_per_graph < s->opt.check_every) s->opt.steps_per_graph = s->opt.check_every;
    s->opt.steps_per_graph -= s->opt.steps_per_graph % s->opt.check_every;
    const GridDims &g = s->g;
    if (g.N0 < 2 * NRPYELL_NG || g.N1 < 2 * NRPYELL_NG || g.N2 < 2 || (g.N2 & 1)) {
      set_last_error("grid too small for the fused solver (need Nxx0,Nxx1 >= 10 and even Nxx2)");
      return NRPYELL_EINVAL;
    }
    if ((unsigned long)3 * g.ntot * sizeof(double) >= (1ul << 32)) {
      // the fused kernels address each array with 32-bit byte offsets
      set_last_error("grid too large for the fused solver: each array must stay under 4 GiB");
      return NRPYELL_EINVAL;
    }
    std::string err;
    if (ghostmap_verify(g, &gd->bcstruct, &err) != NRPYELL_OK) {
      set_last_error(err);
      return NRPYELL_EGHOSTMAP;
    }
    if (s->opt.outer_bc != NRPYELL_OUTER_BC_RADIATION && s->opt.outer_bc != NRPYELL_OUTER_BC_SOMMERFELD6) {
      set_last_error("outer_bc must be NRPYELL_OUTER_BC_RADIATION or NRPYELL_OUTER_BC_SOMMERFELD6");
      return NRPYELL_EINVAL;
    }
    int tx = s->opt.tile_i0 > 0 ? s->opt.tile_i0 : 32;
    if (tx != 16 && tx != 32 && tx != 64) {
      set_last_error("tile_i0 must be 16, 32 or 64");
      return NRPYELL_EINVAL;
    }
    const int pl = s->opt.planes > 0 ? s->opt.planes : (tx >= 32 ? 2 : 1);
    if (pl != 1 && !(pl == 2 && tx >= 32)) {
      set_last_error("planes must be 1, or 2 with tile_i0 32 or 64");
      return NRPYELL_EINVAL;
    }
    s->tx = tx;
    s->pl = pl;
    {
      // two substeps per launch: FP64 steps with the radiation BC; the BC blocks' reflected rows
      // and the phi windows need Nxx1 >= 10 (checked above) and Nxx2 >= 6
      const bool can = s->opt.outer_bc == NRPYELL_OUTER_BC_RADIATION && g.N2 >= 6;
      const int want = s->opt.substeps_per_launch;
      if (want != 0 && want != 1 && want != 2) {
        set_last_error("substeps_per_launch must be 0, 1 or 2");
        return NRPYELL_EINVAL;
      }
      if (want == 2 && !can) {
        set_last_error("two substeps per launch need the radiation outer BC and Nxx2 >= 6");
        return NRPYELL_EINVAL;
      }
      s->pair_path = want == 2 || (want == 0 && can);
    }
    const int ty = NRPYELL_BLOCK / tx;
    s->ntile0 = (g.N0 + tx - 1) / tx;
    s->ntile1 = (g.N1 + ty - 1) / ty;
    {
      // the kernels split block indices with div_by: exact while index x divisor <= 2^32
      // (the two-substep kernel's BC blocks cover fewer rows: NRPYELL_PAIR_BC_ROWS >= 8)
      const double lim = 4294967296.0, nbc = (double)((g.N1 + 7) / 8);
      const double nsub = (double)s->ntile0 * s->ntile1 * (g.N2 / pl);
      if (nsub * (g.N2 / pl) > lim || (double)s->ntile0 * s->ntile1 * s->ntile0 > lim || nbc * nbc * g.N2 > lim) {
        set_last_error("grid too large for the fused solver's block indexing");
        return NRPYELL_EINVAL;
      }
    }
    pair_block_order(s.get());
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    const long n2 = (long)g.G0 * g.G1;
    s->coef64 = dalloc<double>(5 * n2);
    std::vector<double> geo((size_t)6 * g.G1 * 4);
    host_outer_geometry(&gd->params, gd->xx, geo.data());
    s->geo64 = upload(geo.data(), geo.size());
    std::vector<float> geof(geo.begin(), geo.end());
    s->geo32 = upload(geof.data(), geof.size());
    std::vector<double> w(n2);
    host_l2_weights(&gd->params, gd->xx, s->opt.integration_radius, w.data(), &s->volume);
    s->w2d = upload(w.data(), w.size());
    // tables with the wave speed folded in (auxevol and rfm are read here, once per solver)
    s->kc64 = dalloc<double>(6 * n2);
    s->kc32 = dalloc<float>(6 * n2);
    s->src64 = dalloc<double>(2 * g.ntot);
    s->bad_host = halloc_mapped<int>(1, &s->bad);
    s->refresh_tables(s->stream);
    HIPCHK(hipStreamSynchronize(s->stream));
    if (*(volatile int *)s->bad_host) {
      set_last_error("the wave speed gridfunction varies along xx2; the solver needs c = c(xx0, xx1) as "
                     "wavespeed_gf_all_points sets it");
      return NRPYELL_EINVAL;
    }
    const long nblk_res = (long)((g.N0 + 63) / 64) * ((g.N1 + 3) / 4) * g.N2;
    const long nblk_sub = (long)s->ntile0 * s->ntile1 * (g.N2 / pl);
    s->partial = dalloc<double>(nblk_res > nblk_sub ? nblk_res : nblk_sub);
    s->l2_out = dalloc<double>(1);
    s->dstate = dalloc<SolverDeviceState>(1);
    HIPCHK(hipHostMalloc((void **)&s->hstate, sizeof(SolverDeviceState), hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    reset_state(s.get());
    *out = s.release();
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_solver_destroy(nrpyell_solver *s) {
  return api_call([&]() -> int {
    if (!s) return NRPYELL_OK;
    (void)hipStreamSynchronize(s->stream);
    delete s;
    return NRPYELL_OK;
  });
}

static void ensure_fp32(nrpyell_solver *s) {
  if (s->y32) return;
  const long nt = s->g.ntot;
  s->y32 = dalloc<float>(2 * nt);
  s->ko32 = dalloc<float>(2 * nt);
  s->ke32 = dalloc<float>(2 * nt);
  s->rt32 = dalloc<float>(2 * nt);
  s->src32 = dalloc<float>(2 * nt);
}

// Capture `nsteps` RK4 steps with a residual check every check_every steps as one graph.
static hipGraphExec_t capture(nrpyell_solver *s, bool fp32, int nsteps, double target) {
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 1; k <= nsteps; k++) {
    if (fp32) {
      // FP32 bulk phase: the residual is evaluated in FP64 from the FP32 state every check_every
      // steps; on the pair path the state after step k is in ycur32(k & 1) (nsteps is even)
      if (s->pair_path) s->step32_pair(s->stream, (k - 1) & 1);
      else s->step32(s->stream);
      if (k % s->opt.check_every == 0)
        s->check(s->stream, true, s->opt.check_every, target, s->pair_path ? s->ycur32(k & 1) : s->y32);
    } else if (s->pair_path) {
      // FP64, two substeps per launch: y alternates between the two arrays (nsteps is even)
      s->step64_pair(s->stream, (k - 1) & 1, true);
    } else {
      // FP64: every step's stage-1 kernel checks the residual of the state it starts from
      s->step64(s->stream, true);
    }
  }
  HIPCHK(hipStreamEndCapture(s->stream, &graph));
  hipGraphExec_t exec;
  HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HIPCHK(hipGraphDestroy(graph));
  return exec;
}

extern "C" int nrpyell_solver_steps(nrpyell_solver *s, int n) {
  return api_call([&]() -> int {
    reset_state(s);
    if (s->pair_path) {
      for (int k = 0; k < n; k++) s->step64_pair(s->stream, k & 1, false);
      s->finish(s->stream, n & 1);
    } else {
      for (int k = 0; k < n; k++) s->step64(s->stream);
      s->finish(s->stream, 0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_solver_residual_l2(nrpyell_solver *s, double *log10_l2) {
  return api_call([&]() -> int {
    const long npart = launch_residual_l2(s->g, s->gd->gridfuncs.y_n_gfs, false, s->gd->gridfuncs.auxevol_gfs,
                                          s->coef64, s->w2d, s->partial, nullptr, nullptr, s->stream);
    launch_l2_finalize(s->partial, npart, s->volume, nullptr, 0, 0.0, s->l2_out, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(log10_l2, s->l2_out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_solver_solve(nrpyell_solver *s, nrpyell_solve_result *res) {
  return api_call([&]() -> int {
    memset(res, 0, sizeof(*res));
    const auto t0 = std::chrono::steady_clock::now();
    const long nt = s->g.ntot;
    MoL_gridfunctions_struct &f = s->gd->gridfuncs;
    reset_state(s);
    // residual of the initial guess (counts no steps)
    s->check(s->stream, false, 0, s->opt.log10_target);
    read_state(s);
    res->log10_l2_initial = s->hstate->last_log10;
    int steps = 0;
    if (s->hstate->done == 0 && s->opt.precision == 32) {
      ensure_fp32(s);
      launch_convert_d2f(2 * nt, f.y_n_gfs, s->y32, s->stream);
      launch_convert_d2f(2 * nt, s->src64, s->src32, s->stream);
      const nrpyell_solver::GraphKey k32 = s->graph_key(true, s->opt.fp32_switch_log10);
      if (!s->graph32 || !s->key32.same(k32)) {
        if (s->graph32) (void)hipGraphExecDestroy(s->graph32);
        s->graph32 = nullptr;
        s->graph32 = capture(s, true, s->graph_steps32(), s->opt.fp32_switch_log10);
        s->key32 = k32;
      }
      reset_state(s);
      // The relaxation residual oscillates, so the FP32 phase ends on the lowest value seen so
      // far: at the switch target, after fp32_plateau_steps steps without a new lowest value
      // (FP32 rounding floor), or when it climbs 2 decades above the lowest (diverging).
      double best = res->log10_l2_initial;
      int since_best = 0, seen = 0;
      while (true) {
        HIPCHK(hipGraphLaunch(s->graph32, s->stream));
        read_state(s);
        const SolverDeviceState &h = *s->hstate;
        if (h.done || h.steps >= s->opt.max_steps) break;
        if (h.nhist > seen) {
          for (int i = seen; i < h.nhist; i++) {
            if (h.hist_log10[i] < best - 0.01) {
              best = h.hist_log10[i];
              since_best = 0;
            } else {
              since_best += s->opt.check_every;
            }
          }
          seen = h.nhist;
        } else if (h.last_log10 < best - 0.01) {  // history full: the graph's last check
          best = h.last_log10;
          since_best = 0;
        } else {
          since_best += s->graph_steps32();
        }
        if (since_best >= s->opt.fp32_plateau_steps) break;
        if (h.last_log10 > best + 2.0) break;
      }
      steps = s->hstate->steps;
      res->fp32_steps = steps;
      std::vector<int> h32s(s->hstate->hist_step, s->hstate->hist_step + s->hstate->nhist);
      std::vector<double> h32l(s->hstate->hist_log10, s->hstate->hist_log10 + s->hstate->nhist);
      // a check that ended the phase leaves the state of its step count (pair path: in ycur32 of
      // that parity; whole graphs run an even number of steps)
      launch_convert_f2d(2 * nt, s->pair_path ? s->ycur32(steps & 1) : s->y32, f.y_n_gfs, s->stream);
      reset_state(s);
      s->check(s->stream, false, 0, s->opt.log10_target);
      read_state(s);
      // the FP64 phase records its own check at the switch step
      while (!h32s.empty() && h32s.back() >= steps) {
        h32s.pop_back();
        h32l.pop_back();
      }
      s->hist32_step = std::move(h32s);
      s->hist32_log10 = std::move(h32l);
    }
    const std::vector<int> keep_s = s->hist32_step;
    const std::vector<double> keep_l = s->hist32_log10;
    int at = 0;  // the array holding the final state (pair path)
    if (s->hstate->done == 0) {
      const nrpyell_solver::GraphKey k64 = s->graph_key(false, s->opt.log10_target);
      if (!s->graph64 || !s->key64.same(k64)) {
        if (s->graph64) (void)hipGraphExecDestroy(s->graph64);
        s->graph64 = nullptr;
        s->graph64 = capture(s, false, s->graph_steps64(), s->opt.log10_target);
        s->key64 = k64;
      }
      reset_state(s);
      // every step's stage 1 checks the state it starts from: a converged solve took `steps`
      // steps (the count when its check passed), an unconverged one all `nchecks` launched.
      // Under the thorn's rule the device itself ends the solve after step max_steps + 1.
      const bool thorn = s->opt.stop_rule == NRPYELL_STOP_ET_THORN;
      while (true) {
        HIPCHK(hipGraphLaunch(s->graph64, s->stream));
        read_state(s);
        if (s->hstate->done || (!thorn && steps + s->hstate->nchecks >= s->opt.max_steps)) break;
      }
      steps += s->hstate->done ? s->hstate->steps : s->hstate->nchecks;
      // pair path: a check that ended the solve left its state in the array of its parity;
      // otherwise every graph ran an even number of steps
      if (s->pair_path) at = s->hstate->done ? (s->hstate->steps & 1) : 0;
    }
    s->finish(s->stream, at);
    HIPCHK(hipStreamSynchronize(s->stream));
    const auto t1 = std::chrono::steady_clock::now();
    s->hist32_step = keep_s;
    s->hist32_log10 = keep_l;
    s->fp32_offset = res->fp32_steps;
    res->steps = steps;
    res->log10_l2 = s->hstate->last_log10;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    res->converged = s->hstate->done == 1;
    res->status = s->hstate->done == 4 ? NRPYELL_STATUS_MAX_STEPS : s->hstate->done;
    if (s->hstate->done == 2) { set_last_error("residual became NaN"); return NRPYELL_ENAN; }
    if (s->hstate->done == 3) {
      set_last_error("residual diverged (4 decades above its lowest value)");
      return NRPYELL_ENOTCONVERGED;
    }
    if (!res->converged) { set_last_error("max_steps reached before the residual target"); return NRPYELL_ENOTCONVERGED; }
    return NRPYELL_OK;
  });
}

// Exactly n RK4 steps of the production FP64 path (graph replay, stage-1 residual check every
// step against an unreachable target so nothing stops early), bracketed by HIP events on the
// solver's stream. Returns the elapsed milliseconds and the final log10 L2.
extern "C" int nrpyell_solver_timed_steps(nrpyell_solver *s, int n, double *ms_out, double *log10_out) {
  return api_call([&]() -> int {
    const double keep = s->opt.log10_target;
    s->opt.log10_target = -1e300;
    const int per = s->graph_steps64();
    hipGraphExec_t g = capture(s, false, per, s->opt.log10_target);
    s->opt.log10_target = keep;
    hipGraphExec_t tail = nullptr;
    if (n % per) {
      s->opt.log10_target = -1e300;
      tail = capture(s, false, n % per, s->opt.log10_target);
      s->opt.log10_target = keep;
    }
    reset_state(s);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s->stream));
    for (int k = 0; k < n / per; k++) HIPCHK(hipGraphLaunch(g, s->stream));
    if (tail) HIPCHK(hipGraphLaunch(tail, s->stream));
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    s->finish(s->stream, s->pair_path ? (n % per) & 1 : 0);
    read_state(s);
    *ms_out = ms;
    if (log10_out) *log10_out = s->hstate->last_log10;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGraphExecDestroy(g));
    if (tail) HIPCHK(hipGraphExecDestroy(tail));
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_set_eta_damping(griddata_struct *gd, double eta) {
  return api_call([&]() -> int {
    if (!gd || !gd->gridfuncs.y_n_gfs || !(eta > 0.0)) {
      set_last_error("nrpyell_set_eta_damping: null griddata or eta <= 0");
      return NRPYELL_EINVAL;
    }
    const long n = dims_from_params(&gd->params).ntot;
    launch_shift_v(n, eta - gd->params.eta_damping, gd->gridfuncs.y_n_gfs, default_stream());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    gd->params.eta_damping = eta;
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_solver_substeps_per_launch(nrpyell_solver *s) { return s ? (s->pair_path ? 2 : 1) : 0; }

extern "C" int nrpyell_solver_history(nrpyell_solver *s, int max, int *steps, double *log10_l2) {
  return api_call([&]() -> int {
    read_state(s);
    int n = 0;
    for (size_t k = 0; k < s->hist32_step.size() && n < max; k++, n++) {
      steps[n] = s->hist32_step[k];
      log10_l2[n] = s->hist32_log10[k];
    }
    for (int k = 0; k < s->hstate->nhist && n < max; k++, n++) {
      steps[n] = s->fp32_offset + s->hstate->hist_step[k];
      log10_l2[n] = s->hstate->hist_log10[k];
    }
    return n;
  });
}

extern "C" int nrpyell_solver_time_substep(nrpyell_solver *s, int stage, int n, double *us_per_launch) {
  return api_call([&]() -> int {
    MoL_gridfunctions_struct &f = s->gd->gridfuncs;
    reset_state(s);
    if (stage == 5 || stage == 6) {
      // the two-substep kernels: 5 = stages 1+2 with the fused residual check, 6 = stages 3+4;
      // their outputs (k_odd, the running total, k_even) are scratch, y_n is only read
      if (!s->pair_path) {
        set_last_error("this solver runs one substep per launch");
        return NRPYELL_EINVAL;
      }
      SubstepArgs<double> a = s->args<double>(1, f.y_n_gfs, f.k_odd_gfs, nullptr, f.y_nplus1_running_total_gfs,
                                              s->src64, s->kc64, s->geo64);
      a.in = a.yn = f.y_n_gfs;
      a.out = f.k_odd_gfs;
      s->pair_layout(a);
      a.l2_partial = s->partial;
      a.state = s->dstate;
      a.l2_volume = s->volume;
      a.l2_target = -1e300;
      if (stage == 6) {
        a.in = f.k_odd_gfs;
        a.out = f.k_even_gfs;
        a.l2_partial = nullptr;  // time the steady-state launch, not the block-0 reduction
      }
      const int first = stage == 5 ? 1 : 3;
      for (int k = 0; k < 3; k++) launch_substep_pair(a, first, s->stream, &s->order);
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipEventRecord(e0, s->stream));
      for (int k = 0; k < n; k++) launch_substep_pair(a, first, s->stream, &s->order);
      HIPCHK(hipEventRecord(e1, s->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      *us_per_launch = 1000.0 * ms / n;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return NRPYELL_OK;
    }
    // stage 0 = stage 1 with the fused residual-L2 check (the form the solver runs)
    const bool fuse = stage == 0;
    if (fuse) stage = 1;
    // the kernels write y_n only in stage 4; time on scratch copies so the state is untouched
    SubstepArgs<double> a =
        s->args<double>(stage, f.y_n_gfs, f.k_odd_gfs, f.k_even_gfs, f.y_nplus1_running_total_gfs, s->src64,
                        s->kc64, s->geo64);
    if (fuse) {
      a.l2_partial = s->partial;
      a.state = s->dstate;
      a.l2_volume = s->volume;
      a.l2_target = -1e300;  // never "converges": every launch does the full work
    }
    SubstepArgs<double> b = a;
    double *tmp = nullptr;
    if (stage == 4) {
      tmp = dalloc<double>(2 * s->g.ntot);
      HIPCHK(hipMemcpyAsync(tmp, f.y_n_gfs, 2 * s->g.ntot * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
      b.out = tmp;
      b.yn = tmp;
    }
    for (int k = 0; k < 3; k++) launch_substep<double>(b, stage, s->stream);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s->stream));
    for (int k = 0; k < n; k++) launch_substep<double>(b, stage, s->stream);
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *us_per_launch = 1000.0 * ms / n;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (tmp) (void)hipFree(tmp);
    return NRPYELL_OK;
  });
}

// ---------------------------------------------------------------------------------------
// MoL_step_forward_in_time: one RK4 step through a solver cached per griddata.
// ---------------------------------------------------------------------------------------
// The solver reads, at creation, the grid (dimensions, SinhSymTP parameters and the host xx
// arrays: outer-BC geometry, ghost map check) and the buffer addresses; an entry is reused only
// while all of them are unchanged. The rfm and auxevol arrays are read again on every call, as
// the reference's MoL step reads them, so rewriting them in place (new punctures through
// auxevol_gfs_all_points, or any other writer) takes effect at the next step. The free
// functions drop the entries that point at the memory they release.
struct StepCacheEntry {
  nrpyell_solver *s;
  MoL_gridfunctions_struct gf;
  const void *rfm0;
  const innerpt_bc_struct *inner;
  int ninner;
  std::vector<double> grid;
};
static std::map<const griddata_struct *, StepCacheEntry> g_step_cache;
// Guards g_step_cache: MoL_step_forward_in_time holds it for its whole call, the functions that
// free griddata memory while they evict. (Not g_mu: the step creates solvers, which take g_mu.)
static std::mutex g_step_mu;

static std::vector<double> step_grid_key(const griddata_struct *gd) {
  const paramstruct &p = gd->params;
  std::vector<double> k = {(double)p.Nxx0, (double)p.Nxx1, (double)p.Nxx2, p.AMAX, p.SINHWAA, p.bScale,
                           p.dxx0, p.dxx1, p.dxx2, p.xxmin0, p.xxmin1, p.xxmin2};
  const int G[3] = {p.Nxx_plus_2NGHOSTS0, p.Nxx_plus_2NGHOSTS1, p.Nxx_plus_2NGHOSTS2};
  for (int d = 0; d < 3; d++) k.insert(k.end(), gd->xx[d], gd->xx[d] + G[d]);
  return k;
}

template <typename Pred>
static void step_cache_evict(Pred pred) {
  std::lock_guard<std::mutex> lk(g_step_mu);
  for (auto it = g_step_cache.begin(); it != g_step_cache.end();) {
    if (pred(it->second)) {
      nrpyell_solver_destroy(it->second.s);
      it = g_step_cache.erase(it);
    } else {
      ++it;
    }
  }
}

static void step_cache_forget_buffer(const void *p) {
  if (!p) return;
  step_cache_evict([p](const StepCacheEntry &e) {
    return e.gf.y_n_gfs == p || e.gf.k_odd_gfs == p || e.gf.k_even_gfs == p ||
           e.gf.y_nplus1_running_total_gfs == p || e.gf.auxevol_gfs == p || e.rfm0 == p || e.inner == p;
  });
}

extern "C" void MoL_step_forward_in_time(griddata_struct *griddata, const REAL dt) {
  ref_call("MoL_step_forward_in_time", [&]() {
    std::lock_guard<std::mutex> lk(g_step_mu);
    nrpyell_solver *s = nullptr;
    const std::vector<double> key = step_grid_key(griddata);
    const MoL_gridfunctions_struct &gf = griddata->gridfuncs;
    {
      auto it = g_step_cache.find(griddata);
      if (it != g_step_cache.end()) {
        const StepCacheEntry &e = it->second;
        if (e.gf.y_n_gfs == gf.y_n_gfs && e.gf.k_odd_gfs == gf.k_odd_gfs && e.gf.k_even_gfs == gf.k_even_gfs &&
            e.gf.y_nplus1_running_total_gfs == gf.y_nplus1_running_total_gfs && e.gf.auxevol_gfs == gf.auxevol_gfs &&
            e.rfm0 == griddata->rfmstruct.f0_of_xx0 && e.inner == griddata->bcstruct.inner_bc_array &&
            e.ninner == griddata->bcstruct.bc_info.num_inner_boundary_points && e.grid == key) {
          s = e.s;
        } else {
          nrpyell_solver_destroy(e.s);
          g_step_cache.erase(it);
        }
      }
      if (!s) {
        nrpyell_solve_options o;
        nrpyell_default_options(&o);
        if (nrpyell_solver_create(griddata, dt, &o, &s) != NRPYELL_OK) fatal("MoL_step_forward_in_time", g_last_error);
        g_step_cache[griddata] = {s, gf, griddata->rfmstruct.f0_of_xx0, griddata->bcstruct.inner_bc_array,
                                  griddata->bcstruct.bc_info.num_inner_boundary_points, key};
      }
    }
    s->dt = dt;
    // one graph per solver: table refresh (the rfm and auxevol arrays are re-read every call),
    // the RK4 step (two launches of two substeps where the grid allows, then y back to
    // y_n_gfs), the phi ghost planes; recaptured when dt, eta or the outer wave speeds change
    const nrpyell_solver::GraphKey k = s->graph_key(false, 0.0);
    if (!s->graph_step || !s->key_step.same(k)) {
      if (s->graph_step) (void)hipGraphExecDestroy(s->graph_step);
      s->graph_step = nullptr;
      hipGraph_t gr;
      HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      s->refresh_tables(s->stream, &s->dstate->done);
      if (s->pair_path) {
        s->step64_pair(s->stream, 0, false);
        s->finish(s->stream, 1);
      } else {
        s->step64(s->stream);
        s->finish(s->stream, 0);
      }
      HIPCHK(hipStreamEndCapture(s->stream, &gr));
      HIPCHK(hipGraphInstantiate(&s->graph_step, gr, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(gr));
      s->key_step = k;
    }
    HIPCHK(hipGraphLaunch(s->graph_step, s->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    if (*(volatile int *)s->bad_host)
      fatal("MoL_step_forward_in_time",
            "the wave speed gridfunction varies along xx2; the fused step needs c = c(xx0, xx1) as "
            "wavespeed_gf_all_points sets it");
  });
}

// ---------------------------------------------------------------------------------------
// Batched solver: many configurations (each its own grid shape) relaxed together, one
// launch per RK4 substep per tile width, each configuration stopping on its own residual.
// ---------------------------------------------------------------------------------------
__global__ void gather_state_kernel(SolverDeviceState *const *states, int n, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const SolverDeviceState *s = states[k];
  out[4 * k + 0] = (double)s->done;
  out[4 * k + 1] = (double)s->steps;
  out[4 * k + 2] = s->last_log10;
  out[4 * k + 3] = (double)s->nchecks;
}

struct nrpyell_batch {
  int n = 0;
  std::vector<nrpyell_solver *> solvers;
  nrpyell_solve_options opt{};
  hipStream_t stream = nullptr;
  struct Group {
    int tx = 32, pl = 1;
    std::vector<int> members;
    SubstepArgs<double> *d_args[5] = {};   // one substep per launch: per stage 1..4
    SubstepArgs<double> *d_pair[4] = {};   // two per launch: [2 j + l], l = 0 (stages 1+2), 1 (3+4), y in ycur(j)
    int *d_block_cfg = nullptr, *d_block_first = nullptr;  // sized for every member
    int nblocks = 0;                                       // blocks of the members in the graph
    hipStream_t side = nullptr;                            // the group's branch while capturing
    hipEvent_t joined = nullptr;
  };
  std::vector<Group> groups;
  bool pair = false;  // every member runs two substeps per launch (as its single solver would)
  SolverDeviceState **d_states = nullptr;
  double *d_heads = nullptr, *h_heads = nullptr;
  hipGraphExec_t graph = nullptr;
  hipEvent_t fork = nullptr;
  int in_graph = 0;  // configurations the current graph launches blocks for
  ~nrpyell_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (auto &g : groups) {
      for (auto *p : g.d_args) (void)hipFree(p);
      for (auto *p : g.d_pair) (void)hipFree(p);
      (void)hipFree(g.d_block_cfg);
      (void)hipFree(g.d_block_first);
      if (g.side) (void)hipStreamDestroy(g.side);
      if (g.joined) (void)hipEventDestroy(g.joined);
    }
    if (fork) (void)hipEventDestroy(fork);
    (void)hipFree(d_states);
    (void)hipFree(d_heads);
    if (h_heads) (void)hipHostFree(h_heads);
    for (auto *s : solvers) nrpyell_solver_destroy(s);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // One RK4 step of a group: step k of a graph has y in ycur(k & 1) on the pair path.
  void group_step(const Group &g, int k, hipStream_t st) const {
    if (pair) {
      const int j = k & 1;
      launch_substep_pair_batched(g.d_pair[2 * j], g.d_block_cfg, g.d_block_first, g.nblocks, g.tx, g.pl, 1, st);
      launch_substep_pair_batched(g.d_pair[2 * j + 1], g.d_block_cfg, g.d_block_first, g.nblocks, g.tx, g.pl, 3, st);
      return;
    }
    for (int stage = 1; stage <= 4; stage++)
      launch_substep_batched<double>(g.d_args[stage], g.d_block_cfg, g.d_block_first, g.nblocks, g.tx, g.pl, stage,
                                     st);
  }
  // RK4 steps per graph: even on the pair path, so every graph starts with y in y_n_gfs
  int graph_steps() const { return pair ? opt.steps_per_graph + (opt.steps_per_graph & 1) : opt.steps_per_graph; }
  // blocks of one configuration in a launch
  int blocks_of(const nrpyell_solver *s) const {
    const int rows = pair ? NRPYELL_PAIR_BC_ROWS : NRPYELL_BC_ROWS;
    return s->ntile0 * s->ntile1 * (s->g.N2 / s->pl) + (s->g.N1 + rows - 1) / rows * s->g.N2;
  }
  // Block maps over the configurations with active[k] set, and a graph of steps_per_graph
  // steps over them: converged configurations drop out of the launches.
  void rebuild(const std::vector<char> &active) {
    in_graph = 0;
    for (auto &g : groups) {
      const int m = (int)g.members.size();
      std::vector<int> first(m, 0), bcfg;
      for (int j = 0; j < m; j++) {
        const int k = g.members[j];
        if (!active[k]) continue;
        in_graph++;
        nrpyell_solver *s = solvers[k];
        first[j] = (int)bcfg.size();
        const int nb = blocks_of(s);
        for (int q = 0; q < nb; q++) bcfg.push_back(j);
      }
      if (!g.d_block_cfg) {  // first build: every member is active, the buffers get their full size
        g.d_block_cfg = dalloc<int>(bcfg.size() > 0 ? bcfg.size() : 1);
        g.d_block_first = dalloc<int>(m);
      }
      g.nblocks = (int)bcfg.size();
      if (g.nblocks > 0)
        HIPCHK(hipMemcpyAsync(g.d_block_cfg, bcfg.data(), sizeof(int) * bcfg.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(g.d_block_first, first.data(), sizeof(int) * m, hipMemcpyHostToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    if (graph) HIPCHK(hipGraphExecDestroy(graph));
    hipGraph_t gr;
    int nlive = 0;
    for (auto &g : groups) nlive += g.nblocks > 0;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    if (nlive <= 1) {
      for (int k = 0; k < graph_steps(); k++)
        for (auto &g : groups)
          if (g.nblocks > 0) group_step(g, k, stream);
    } else {
      // the groups' configurations are independent: one branch per group (its steps in order,
      // on a stream forked from the capture and joined at the end), so the GPU can overlap the
      // groups' launches instead of running them one after another in every stage
      if (!fork) HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
      HIPCHK(hipEventRecord(fork, stream));
      for (auto &g : groups) {
        if (g.nblocks == 0) continue;
        if (!g.side) HIPCHK(hipStreamCreateWithFlags(&g.side, hipStreamNonBlocking));
        if (!g.joined) HIPCHK(hipEventCreateWithFlags(&g.joined, hipEventDisableTiming));
        HIPCHK(hipStreamWaitEvent(g.side, fork, 0));
        for (int k = 0; k < graph_steps(); k++) group_step(g, k, g.side);
        HIPCHK(hipEventRecord(g.joined, g.side));
        HIPCHK(hipStreamWaitEvent(stream, g.joined, 0));
      }
    }
    HIPCHK(hipStreamEndCapture(stream, &gr));
    HIPCHK(hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(gr));
  }
  void read_heads() {
    hipLaunchKernelGGL(gather_state_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, d_states, n, d_heads);
    HIPCHK(hipMemcpyAsync(h_heads, d_heads, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
};

extern "C" int nrpyell_batch_create(griddata_struct **gds, const double *dts, const int *tile_i0, const int *planes,
                                    int n, const nrpyell_solve_options *opt, nrpyell_batch **out) {
  return api_call([&]() -> int {
    if (n <= 0) { set_last_error("empty batch"); return NRPYELL_EINVAL; }
    auto b = std::make_unique<nrpyell_batch>();
    b->n = n;
    if (opt) b->opt = *opt; else nrpyell_default_options(&b->opt);
    b->opt.precision = 64;
    if (b->opt.stop_rule != NRPYELL_STOP_EVERY_STEP) {
      set_last_error("the batched solver stops every configuration by NRPYELL_STOP_EVERY_STEP");
      return NRPYELL_EINVAL;
    }
    for (int k = 0; k < n; k++) {
      nrpyell_solve_options o = b->opt;
      o.tile_i0 = tile_i0 ? tile_i0[k] : b->opt.tile_i0;
      o.planes = planes ? planes[k] : b->opt.planes;
      nrpyell_solver *s = nullptr;
      const int rc = nrpyell_solver_create(gds[k], dts[k], &o, &s);
      if (rc != NRPYELL_OK) return rc;
      b->solvers.push_back(s);
    }
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->pair = true;
    for (auto *s : b->solvers) b->pair = b->pair && s->pair_path;
    std::map<std::pair<int, int>, std::vector<int>> by_tile;
    for (int k = 0; k < n; k++) by_tile[{b->solvers[k]->tx, b->solvers[k]->pl}].push_back(k);
    for (auto &kv : by_tile) {
      nrpyell_batch::Group g;
      g.tx = kv.first.first;
      g.pl = kv.first.second;
      g.members = kv.second;
      const int m = (int)g.members.size();
      for (int stage = 1; stage <= 4; stage++) {
        std::vector<SubstepArgs<double>> args(m);
        for (int j = 0; j < m; j++) {
          nrpyell_solver *s = b->solvers[g.members[j]];
          MoL_gridfunctions_struct &f = s->gd->gridfuncs;
          SubstepArgs<double> a = s->args<double>(stage, f.y_n_gfs, f.k_odd_gfs, f.k_even_gfs,
                                                  f.y_nplus1_running_total_gfs, s->src64, s->kc64, s->geo64);
          if (stage <= 2) {
            a.l2_partial = s->partial;
            a.state = s->dstate;
            a.l2_volume = s->volume;
            a.l2_target = b->opt.log10_target;
          }
          args[j] = a;
        }
        g.d_args[stage] = upload(args.data(), args.size());
      }
      if (b->pair) {
        // as step64_pair with the fused check: launch 1 writes the residual partials of y_n,
        // launch 2's block 0 reduces them
        for (int j = 0; j < 2; j++)
          for (int l = 0; l < 2; l++) {
            std::vector<SubstepArgs<double>> args(m);
            for (int i = 0; i < m; i++) {
              nrpyell_solver *s = b->solvers[g.members[i]];
              MoL_gridfunctions_struct &f = s->gd->gridfuncs;
              SubstepArgs<double> a = s->args<double>(1, s->ycur(j), f.k_odd_gfs, nullptr,
                                                      f.y_nplus1_running_total_gfs, s->src64, s->kc64, s->geo64);
              a.in = l == 0 ? s->ycur(j) : f.k_odd_gfs;
              a.yn = s->ycur(j);
              a.out = l == 0 ? f.k_odd_gfs : s->ycur(1 - j);
              s->pair_layout(a);
              a.l2_partial = s->partial;
              a.state = s->dstate;
              a.l2_volume = s->volume;
              a.l2_target = b->opt.log10_target;
              args[i] = a;
            }
            g.d_pair[2 * j + l] = upload(args.data(), args.size());
          }
      }
      b->groups.push_back(g);
    }
    std::vector<SolverDeviceState *> st(n);
    for (int k = 0; k < n; k++) st[k] = b->solvers[k]->dstate;
    b->d_states = upload(st.data(), st.size());
    b->d_heads = dalloc<double>(4 * n);
    HIPCHK(hipHostMalloc((void **)&b->h_heads, sizeof(double) * 4 * n, hipHostMallocDefault));
    b->rebuild(std::vector<char>(n, 1));
    *out = b.release();
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_batch_solve(nrpyell_batch *b, nrpyell_solve_result *res, double *seconds) {
  return api_call([&]() -> int {
    const auto t0 = std::chrono::steady_clock::now();
    for (auto *s : b->solvers) reset_state(s);
    if (b->in_graph != b->n) b->rebuild(std::vector<char>(b->n, 1));
    int launched = 0;
    std::vector<char> active(b->n, 1);
    while (true) {
      HIPCHK(hipGraphLaunch(b->graph, b->stream));
      launched += b->graph_steps();
      b->read_heads();
      int n_active = 0;
      for (int k = 0; k < b->n; k++) {
        active[k] = b->h_heads[4 * k] == 0.0;
        n_active += active[k];
      }
      if (n_active == 0 || launched >= b->opt.max_steps) break;
      // compact when a quarter of the launched configurations have finished
      if (4 * n_active <= 3 * b->in_graph) b->rebuild(active);
    }
    for (int k = 0; k < b->n; k++) {
      nrpyell_solver *s = b->solvers[k];
      // pair path: a configuration stopped by its check holds the state it checked in the array
      // of that step count's parity; whole graphs run an even number of steps
      s->finish(b->stream, b->pair && b->h_heads[4 * k] != 0.0 ? ((int)b->h_heads[4 * k + 1] & 1) : 0);
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    const auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    int rc = NRPYELL_OK;
    for (int k = 0; k < b->n; k++) {
      nrpyell_solve_result &r = res[k];
      memset(&r, 0, sizeof(r));
      r.converged = b->h_heads[4 * k] == 1.0;
      r.status = (int)b->h_heads[4 * k];
      // converged, NaN or diverged: the steps when its check decided; running: all launched
      r.steps = b->h_heads[4 * k] != 0.0 ? (int)b->h_heads[4 * k + 1] : launched;
      r.log10_l2 = b->h_heads[4 * k + 2];
      r.seconds = *seconds;
      if (b->h_heads[4 * k] == 2.0) rc = NRPYELL_ENAN;
      else if (!r.converged && rc == NRPYELL_OK) rc = NRPYELL_ENOTCONVERGED;
    }
    if (rc == NRPYELL_ENAN) set_last_error("a configuration's residual became NaN");
    if (rc == NRPYELL_ENOTCONVERGED) set_last_error("some configurations diverged or reached max_steps");
    return rc;
  });
}

extern "C" int nrpyell_batch_destroy(nrpyell_batch *b) {
  return api_call([&]() -> int {
    if (!b) return NRPYELL_OK;
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return NRPYELL_OK;
  });
}
